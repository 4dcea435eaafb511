// eb_sim.hip -- the driver of the drop evolution (include/eb_sim.h): initial shapes,
// redistribution to equal arclength, the RK2 time loop with the reference's step law and
// point doubling, and the geometry.data format.
//
// Device work: the initial parametrisations (one thread per point), CLUSTPARAM's Newton
// solves (one workgroup per body, one thread per point), EXTRACTSOL + GETFOURIER after each
// step and DOUBLE's interpolation (one workgroup per body, FFT in LDS), and F.  The host keeps
// the per-body scalars (area, arclength) and the step law, as the reference's main program
// does.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/eb_sim.h"
#include "eb_bodies.h"

using namespace eb;

namespace {

constexpr double PI = 3.14159265358979323846;

// INITPARAM (electric.f:431-500): point and first two derivatives at parameter th
__device__ void initparam(double th, double ai, double bi, double thetax, int ncyc, double& xp, double& yp,
                          double& xdot, double& ydot, double& xddot, double& yddot) {
  if (ncyc == 0) {
    double cs = cos(th - thetax), sn = sin(th - thetax);
    const double a2 = ai * ai, b2 = bi * bi;
    const double rnorm = sqrt(b2 * cs * cs + a2 * sn * sn);
    const double radius = ai * bi / rnorm;
    const double rdot = -ai * bi * cs * sn * (-b2 + a2) / (rnorm * rnorm * rnorm);
    const double rddot = -ai * bi *
                         (2.0 * a2 * b2 * cs * cs * sn * sn + a2 * a2 * cs * cs * cs * cs + a2 * b2 - a2 * a2 +
                          b2 * b2 * cs * cs * cs * cs - 2.0 * b2 * b2 * cs * cs) /
                         (rnorm * rnorm * rnorm * rnorm * rnorm);
    xp = radius * cos(th);
    yp = radius * sin(th);
    cs = cos(th);
    sn = sin(th);
    xdot = rdot * cs - radius * sn;
    ydot = rdot * sn + radius * cs;
    xddot = rddot * cs - 2.0 * rdot * sn - radius * cs;
    yddot = rddot * sn + 2.0 * rdot * cs - radius * sn;
  } else if (ncyc < 0) {
    const double R = ai, anu = bi, s2 = sqrt(2.0), sq = sqrt(1.0 + anu * anu);
    const double den = (1.0 - 2.0 * anu * cos(2.0 * th) + anu * anu) * sq;
    const double ddot = 4.0 * anu * sin(2.0 * th) * sq;
    const double dddot = 8.0 * anu * cos(2.0 * th) * sq;
    const double ax = (1.0 - anu * anu) * (1.0 - anu) * R * s2, ay = (1.0 - anu * anu) * (1.0 + anu) * R * s2;
    xp = ax * cos(th) / den;
    yp = ay * sin(th) / den;
    xdot = -ax * sin(th) / den - xp * ddot / den;
    xddot = -xp + ax * sin(th) * ddot / (den * den) - xdot * ddot / den - xp * dddot / den +
            xp * ddot * ddot / (den * den);
    ydot = ay * cos(th) / den - yp * ddot / den;
    yddot = -yp - ay * cos(th) * ddot / (den * den) - ydot * ddot / den - yp * dddot / den +
            yp * ddot * ddot / (den * den);
  } else {
    const double cs = cos(th), sn = sin(th), snn = sin(ncyc * th), csn = cos(ncyc * th);
    xp = ai * cs + bi * csn;
    yp = ai * sn - bi * snn;
    xdot = -ai * sn - ncyc * bi * snn;
    xddot = -ai * cs - (double)(ncyc * ncyc) * bi * csn;
    ydot = ai * cs - ncyc * bi * csn;
    yddot = -ai * sn + (double)(ncyc * ncyc) * bi * snn;
  }
}

struct ShapeView {
  const double *ai, *bi, *thetax, *cx, *cy;
  const int* ncyc;
};

// GEOINIT (beta == NULL: equispaced theta) or GEOSHIFT (theta = beta): points as interleaved
// (x, y), |z_theta| and the curvature of every point
__global__ void k_initshape(int n, const int* __restrict__ body_of, const int* __restrict__ off,
                            const int* __restrict__ nd, ShapeView sv, const double* __restrict__ beta,
                            double2* __restrict__ z, double* __restrict__ dsdth, double* __restrict__ rkappa) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int k = body_of[i];
  const int j = i - off[k];
  const double th = beta ? beta[i] : (2.0 * PI / nd[k]) * (j - 0.0);
  double xp, yp, xd, yd, xdd, ydd;
  initparam(th, sv.ai[k], sv.bi[k], sv.thetax[k], sv.ncyc[k], xp, yp, xd, yd, xdd, ydd);
  z[i] = make_double2(sv.cx[k] + xp, sv.cy[k] + yp);
  const double s = sqrt(xd * xd + yd * yd);
  if (dsdth) dsdth[i] = s;
  if (rkappa) rkappa[i] = (xd * ydd - yd * xdd) / (s * s * s);
}

// FOUREVAL (electric.f:1336-1355): c_0 + sum_{i=1}^{n-1} c_i e^{i i x} + c_{N-i} e^{-i i x}
__device__ double foureval_re(const double2* c, int N, double x) {
  const int n = N / 2;
  double s = c[0].x;
  for (int i = 1; i <= n - 1; ++i) {
    double sn, cs;
    sincos((double)i * x, &sn, &cs);
    const double2 a = c[i], b = c[N - i];
    // Re(a e^{i t}) + Re(b e^{-i t})
    s += (a.x * cs - a.y * sn) + (b.x * cs + b.y * sn);
  }
  return s;
}

// CLUSTPARAM (electric.f:1192-1265) per body: csa = FFT(|z_theta|)/N with the N/2 term zeroed,
// ARCL = 2 pi Re(csa_0), csai = FOURINT(csa) with the constant making s(0) = 0, then Newton on
// s(beta_i) = i ARCL / N for every point (tolerance 1e-14, at most 100 steps; a failure flags
// the body).
__global__ __launch_bounds__(SP_THREADS) void k_clustparam(BodyView B, const double* __restrict__ dsdth,
                                                            double* __restrict__ beta, double* __restrict__ arcl,
                                                            int* __restrict__ fail) {
  extern __shared__ double2 lds[];
  const BodyCtx bc = body_ctx(B);
  const int N = bc.N, o0 = bc.o0, n = N / 2;
  double2* a = lds;
  double2* c = lds + B.nmax;
  double2* ws = lds + 2 * B.nmax;
  for (int i = threadIdx.x; i < N; i += blockDim.x) a[i] = make_double2(dsdth[o0 + i], 0.0);
  __syncthreads();
  fft(a, ws, bc, B, -1);
  const double invn = 1.0 / N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    double2 v = make_double2(a[i].x * invn, a[i].y * invn);
    if (i == n) v = make_double2(0.0, 0.0);  // csa(n+1) = 0
    a[i] = v;
    // FOURINT (electric.f:1307-1331)
    double2 r = make_double2(0, 0);
    if (i >= 1 && i <= n - 1) {
      r = make_double2(v.y / i, -v.x / i);
    } else if (i >= N - n + 1 && i <= N - 1) {
      const double ii = (double)(N - i);
      r = make_double2(-v.y / ii, v.x / ii);
    }
    c[i] = r;
  }
  __syncthreads();
  const double al = 2.0 * PI * a[0].x;
  if (threadIdx.x == 0) {
    arcl[bc.k] = al;
    c[0] = make_double2(0.0, 0.0);
  }
  __syncthreads();
  const double fk0 = foureval_re(c, N, 0.0);
  __syncthreads();
  if (threadIdx.x == 0) c[0] = make_double2(-fk0, 0.0);
  __syncthreads();
  const double c0 = a[0].x;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    // start from the equispaced parameter (the reference starts from the previous point's
    // root; the roots agree to the Newton tolerance)
    double b0 = 2.0 * PI * i / N;
    bool ok = false;
    for (int it = 0; it < 100; ++it) {
      const double fk = foureval_re(c, N, b0) + c0 * b0 - (double)i * al / N;
      const double fpk = foureval_re(a, N, b0);
      const double del = -fk / fpk;
      b0 += del;
      if (fabs(del) < 1e-14) {
        ok = true;
        break;
      }
    }
    if (!ok) atomicExch(&fail[bc.k], 1);
    beta[o0 + i] = b0;
  }
}

// EXTRACTSOL (electric.f:575-631) and GETFOURIER's doubling test (electric.f:531-570) per
// body.  out[k] = (area, arcl, errds, kappa max, idoub)
__global__ __launch_bounds__(SP_THREADS) void k_sim_post(BodyView B, const double2* __restrict__ ybar,
                                                          double* __restrict__ out) {
  extern __shared__ double2 lds[];
  __shared__ double red[SP_THREADS];
  const BodyCtx bc = body_ctx(B);
  const int N = bc.N, o0 = bc.o0, n = N / 2;
  double2* a = lds;
  double2* c = lds + B.nmax;
  double2* ws = lds + 2 * B.nmax;
  double mydbl = 0.0;
  for (int i = threadIdx.x; i < N; i += blockDim.x) a[i] = ybar[o0 + i];
  __syncthreads();
  fft(a, ws, bc, B, -1);
  const double invn = 1.0 / N;
  // GETFOURIER: fmode(i+1) = |za(i+1)| + |za(N-i+1)| over the last ten modes below N/2
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const double2 v = make_double2(a[i].x * invn, a[i].y * invn);
    if (i >= n - 10 && i <= n - 1 && i >= 1) {
      const double2 w = make_double2(a[N - i].x * invn, a[N - i].y * invn);
      if (hypot(v.x, v.y) + hypot(w.x, w.y) > 1e-11) mydbl = 1.0;
    }
    // FDIFFF: the derivative's coefficients (i k / N) F; dz = -z_alpha
    const double kk = fdiff_k(i, N) * invn;
    c[i] = make_double2(kk * a[i].y, -kk * a[i].x);  // -(i k) F / N
  }
  __syncthreads();
  fft(c, ws, bc, B, +1);  // dz
  // za = FDIFFF(dz)
  for (int i = threadIdx.x; i < N; i += blockDim.x) a[i] = c[i];
  __syncthreads();
  fft(a, ws, bc, B, -1);
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const double kk = fdiff_k(i, N) * invn;
    const double2 v = a[i];
    a[i] = make_double2(-kk * v.y, kk * v.x);
  }
  __syncthreads();
  fft(a, ws, bc, B, +1);
  double smin = 1e10, smax = 0.0, kmax = 0.0, sa = 0.0, sl = 0.0;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const double2 d = c[i], dd = a[i];
    const double ds = sqrt(d.x * d.x + d.y * d.y);
    const double dsa = hypot(d.x, d.y);
    smin = fmin(smin, dsa);
    smax = fmax(smax, dsa);
    kmax = fmax(kmax, fabs((d.x * dd.y - d.y * dd.x) / (ds * ds * ds)));
    sa += ybar[o0 + i].y * d.x;
    sl += dsa;
  }
  const double area = fabs(2.0 * PI * block_sum(sa, red) / N);
  const double arcl = fabs(2.0 * PI * block_sum(sl, red) / N);
  red[threadIdx.x] = smin;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  const double mn = red[0];
  __syncthreads();
  red[threadIdx.x] = smax;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  const double mx = red[0];
  __syncthreads();
  red[threadIdx.x] = kmax;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  const double kap = red[0];
  __syncthreads();
  red[threadIdx.x] = mydbl;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double* r = out + 5 * bc.k;
    r[0] = area;
    r[1] = arcl;
    r[2] = (mx - mn) / mx;
    r[3] = kap;
    r[4] = red[0];
  }
}

// DOUBLE (electric.f:721-757) by FINTERC's interpolation for the bodies with dbl[k]; the others
// are copied.  Old body k at ybar[off_old[k]], new at ynew[off_new[k]].
__global__ __launch_bounds__(SP_THREADS) void k_sim_double(BodyView B, const double2* __restrict__ ybar,
                                                            const int* __restrict__ dbl,
                                                            const int* __restrict__ off_new,
                                                            double2* __restrict__ ynew) {
  extern __shared__ double2 lds[];
  const BodyCtx bc = body_ctx(B);
  const int N = bc.N, o0 = bc.o0, o1 = off_new[bc.k];
  if (!dbl[bc.k]) {
    for (int i = threadIdx.x; i < N; i += blockDim.x) ynew[o1 + i] = ybar[o0 + i];
    return;
  }
  double2* a = lds;
  double2* c = lds + B.nmax;
  double2* ws = lds + 2 * B.nmax;
  for (int i = threadIdx.x; i < N; i += blockDim.x) a[i] = ybar[o0 + i];
  __syncthreads();
  fft(a, ws, bc, B, -1);
  const double invn = 1.0 / N;
  const int nyk = N / 2;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    double2 v = make_double2(a[i].x * invn, a[i].y * invn);
    if (i == nyk) v = make_double2(0.0, 0.0);
    a[i] = v;
    const int ks = (i <= nyk) ? i : i - N;
    double sn, cs;
    sincospi((double)ks / (double)N, &sn, &cs);
    c[i] = make_double2(v.x * cs - v.y * sn, v.x * sn + v.y * cs);
  }
  __syncthreads();
  fft(a, ws, bc, B, +1);
  fft(c, ws, bc, B, +1);
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    ynew[o1 + 2 * i] = a[i];
    ynew[o1 + 2 * i + 1] = c[i];
  }
}

__global__ void k_body_of_s(const int* __restrict__ off, int nbody, int* __restrict__ body_of) {
  const int k = blockIdx.x;
  for (int i = off[k] + threadIdx.x; i < off[k + 1]; i += blockDim.x) body_of[i] = k;
}

}  // namespace

// ------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------

struct eb_sim_s {
  eb_sim_params p{};
  hipStream_t st = nullptr;
  int k = 0, nbk = 0, steps = 0, doublings = 0, nplot = 0;
  double time = 0.0, dt = 0.0, tfinal = 0.0, cmin = 0.0, areatot0 = 0.0, areatot = 0.0, erra = 0.0;
  std::vector<int> nd;
  std::vector<double> area0, arcl0, area, arcl, post;
  eb_bodies B = nullptr;
  eb_field F = nullptr;
  double *ybar = nullptr, *y = nullptr, *dy = nullptr, *phi0 = nullptr, *dpost = nullptr;
};

static void sim_free_state(eb_sim S) {
  if (S->F) eb_field_destroy(S->F);
  if (S->B) eb_bodies_destroy(S->B);
  S->F = nullptr;
  S->B = nullptr;
  void* p[] = {S->ybar, S->y, S->dy, S->phi0, S->dpost};
  for (void* q : p)
    if (q) (void)hipFree(q);
  S->ybar = S->y = S->dy = S->phi0 = S->dpost = nullptr;
}

static void sim_free(eb_sim S) {
  if (!S) return;
  if (S->st) (void)hipStreamSynchronize(S->st);
  sim_free_state(S);
  delete S;
}

// Bodies, F and the RK2 buffers for the current nd
static int sim_alloc(eb_sim S) {
  int rc;
  if ((rc = eb_bodies_create(&S->B, S->k, S->nd.data(), (void*)S->st)) != EB_OK) return rc;
  if ((rc = eb_field_create(&S->F, S->B, S->p.cap, S->p.e0, (void*)S->st)) != EB_OK) return rc;
  S->nbk = S->B->n;
  const size_t neq = 2 * (size_t)S->nbk;
  if ((rc = dalloc(&S->ybar, neq)) || (rc = dalloc(&S->y, neq)) || (rc = dalloc(&S->dy, neq)) ||
      (rc = dalloc(&S->phi0, neq)) || (rc = dalloc(&S->dpost, 5 * (size_t)S->k)))
    return rc;
  S->post.assign(5 * (size_t)S->k, 0.0);
  return EB_OK;
}

static int check_params(const eb_sim_params* p) {
  if (!p || p->k < 1 || p->k > EB_SIM_KMAX) return (eb_set_error("eb_sim: K outside 1..%d", EB_SIM_KMAX), EB_ERR_ARG);
  for (int b = 0; b < p->k; ++b)
    if (p->nd[b] < 16 || p->nd[b] > EB_SIM_NDMAX || (p->nd[b] & 1))
      return (eb_set_error("eb_sim: body %d has ND = %d (even, 16..%d)", b, p->nd[b], EB_SIM_NDMAX), EB_ERR_ARG);
  if (p->bondx != 0.0 || p->bondy != 0.0)
    return (eb_set_error("eb_sim: gravity (BOND not 0) runs with a wall in the reference and is not built"), EB_ERR_ARG);
  if (!(p->dt > 0.0)) return (eb_set_error("eb_sim: DT must be positive"), EB_ERR_ARG);
  return EB_OK;
}

static void sim_cmin(eb_sim S) {
  S->cmin = 0.0;
  for (int b = 0; b < S->k; ++b) S->cmin = std::max(S->cmin, S->dt * S->nd[b] / S->arcl0[b]);  // electric.f:101-102
  S->areatot0 = 0.0;
  for (int b = 0; b < S->k; ++b) S->areatot0 += S->area0[b];
}

extern "C" int eb_sim_create(eb_sim* out, const eb_sim_params* p, void* stream) {
  if (!out) return (eb_set_error("eb_sim_create: bad arguments"), EB_ERR_ARG);
  *out = nullptr;
  int rc = check_params(p);
  if (rc != EB_OK) return rc;
  eb_sim S = new eb_sim_s();
  S->p = *p;
  S->st = (hipStream_t)stream;
  S->k = p->k;
  S->nd.assign(p->nd, p->nd + p->k);
  S->dt = p->dt;
  S->tfinal = p->tfinal;
  S->nplot = p->nplot;
  if ((rc = sim_alloc(S)) != EB_OK) {
    sim_free(S);
    return rc;
  }
  // GEOINIT's areas (electric.f:361-366)
  S->area0.resize(S->k);
  for (int b = 0; b < S->k; ++b)
    S->area0[b] = p->ncyc[b] == 0 ? PI * p->ai[b] * p->bi[b] : PI * (p->ai[b] * p->ai[b] - p->ncyc[b] * p->bi[b] * p->bi[b]);
  // shape parameters on the device
  const int k = S->k, n = S->nbk;
  double *d_ai = nullptr, *d_bi = nullptr, *d_th = nullptr, *d_cx = nullptr, *d_cy = nullptr, *dsdth = nullptr,
         *beta = nullptr, *arcl = nullptr;
  int *d_ncyc = nullptr, *body_of = nullptr, *fail = nullptr;
  auto cleanup = [&]() {
    void* q[] = {d_ai, d_bi, d_th, d_cx, d_cy, dsdth, beta, arcl, d_ncyc, body_of, fail};
    for (void* v : q)
      if (v) (void)hipFree(v);
  };
  std::vector<double> ai(p->ai, p->ai + k), bi(p->bi, p->bi + k), th(p->thetax, p->thetax + k), cx(p->cx, p->cx + k),
      cy(p->cy, p->cy + k);
  std::vector<int> ncyc(p->ncyc, p->ncyc + k), zero(k, 0);
  if ((rc = dalloc(&d_ai, k)) || (rc = dalloc(&d_bi, k)) || (rc = dalloc(&d_th, k)) || (rc = dalloc(&d_cx, k)) ||
      (rc = dalloc(&d_cy, k)) || (rc = dalloc(&d_ncyc, k)) || (rc = dalloc(&dsdth, n)) || (rc = dalloc(&beta, n)) ||
      (rc = dalloc(&arcl, k)) || (rc = dalloc(&body_of, n)) || (rc = dalloc(&fail, k))) {
    cleanup();
    sim_free(S);
    return rc;
  }
  if (!upload(d_ai, ai, S->st) || !upload(d_bi, bi, S->st) || !upload(d_th, th, S->st) || !upload(d_cx, cx, S->st) ||
      !upload(d_cy, cy, S->st) || !upload(d_ncyc, ncyc, S->st) || !upload(fail, zero, S->st)) {
    cleanup();
    sim_free(S);
    return (eb_set_error("eb_sim_create: upload failed"), EB_ERR_HIP);
  }
  eb_bodies B = S->B;
  const ShapeView sv{d_ai, d_bi, d_th, d_cx, d_cy, d_ncyc};
  hipLaunchKernelGGL(k_body_of_s, dim3(k), dim3(256), 0, S->st, B->d_off, k, body_of);
  // GEOINIT, CLUSTPARAM, GEOSHIFT (electric.f:74-79), INITCOND
  hipLaunchKernelGGL(k_initshape, dim3((n + 255) / 256), dim3(256), 0, S->st, n, body_of, B->d_off, B->d_nd, sv,
                     (const double*)nullptr, (double2*)S->ybar, dsdth, (double*)nullptr);
  hipLaunchKernelGGL(k_clustparam, dim3(k), dim3(SP_THREADS), lds_bytes(B, 2), S->st, view(B), dsdth, beta, arcl, fail);
  hipLaunchKernelGGL(k_initshape, dim3((n + 255) / 256), dim3(256), 0, S->st, n, body_of, B->d_off, B->d_nd, sv, beta,
                     (double2*)S->ybar, (double*)nullptr, (double*)nullptr);
  S->arcl0.assign(k, 0.0);
  std::vector<int> fl(k, 0);
  const bool ok = hipGetLastError() == hipSuccess &&
                  hipMemcpyAsync(S->arcl0.data(), arcl, sizeof(double) * k, hipMemcpyDeviceToHost, S->st) == hipSuccess &&
                  hipMemcpyAsync(fl.data(), fail, sizeof(int) * k, hipMemcpyDeviceToHost, S->st) == hipSuccess &&
                  hipStreamSynchronize(S->st) == hipSuccess;
  cleanup();
  if (!ok) {
    sim_free(S);
    return (eb_set_error("eb_sim_create: initial shapes failed"), EB_ERR_HIP);
  }
  for (int b = 0; b < k; ++b)
    if (fl[b]) {
      sim_free(S);
      return (eb_set_error("eb_sim_create: CLUSTPARAM's Newton iteration did not converge on body %d", b), EB_ERR_ARG);
    }
  S->area = S->area0;
  S->arcl = S->arcl0;
  sim_cmin(S);
  *out = S;
  return EB_OK;
}

extern "C" int eb_sim_step(eb_sim S, int32_t* done) {
  if (!S) return (eb_set_error("eb_sim_step: bad arguments"), EB_ERR_ARG);
  hipStream_t st = S->st;
  void* s = (void*)st;
  // electric.f:144-149
  S->time += S->dt;
  if (S->time > S->tfinal) {
    S->dt = S->tfinal - S->time + S->dt;
    S->time = S->tfinal;
  }
  const int neq = 2 * S->nbk;
  int rc;
  // RK2 (electric.f:762-786)
  if ((rc = eb_field_rhs(S->F, S->ybar, S->dy)) != EB_OK) return rc;
  if ((rc = eb_rk2_stage1(neq, S->dt, S->ybar, S->dy, S->phi0, S->y, s)) != EB_OK) return rc;
  if ((rc = eb_field_rhs(S->F, S->y, S->dy)) != EB_OK) return rc;
  if ((rc = eb_rk2_stage2(neq, S->dt, S->phi0, S->dy, S->ybar, s)) != EB_OK) return rc;
  // EXTRACTSOL + GETFOURIER of every body
  eb_bodies B = S->B;
  hipLaunchKernelGGL(k_sim_post, dim3(S->k), dim3(SP_THREADS), lds_bytes(B, 2), st, view(B), (const double2*)S->ybar,
                     S->dpost);
  EB_HIP_CHECK(hipGetLastError());
  EB_HIP_CHECK(hipMemcpyAsync(S->post.data(), S->dpost, sizeof(double) * 5 * S->k, hipMemcpyDeviceToHost, st));
  EB_HIP_CHECK(hipStreamSynchronize(st));
  S->areatot = 0.0;
  std::vector<int> dbl(S->k, 0);
  bool icut = false;
  for (int b = 0; b < S->k; ++b) {
    S->area[b] = S->post[5 * b];
    S->arcl[b] = S->post[5 * b + 1];
    S->areatot += S->area[b];
    dbl[b] = S->post[5 * b + 4] != 0.0;
    icut = icut || dbl[b];
  }
  // A body that needs doubling past F's limit ends the run, as the reference STOPs when DOUBLE
  // would pass NDMAX (electric.f:190-199); the step itself has been taken.
  for (int b = 0; b < S->k; ++b)
    if (dbl[b] && 2 * S->nd[b] > EB_SIM_NDMAX) {
      ++S->steps;
      return (eb_set_error("eb_sim_step: body %d needs doubling to %d points, past the limit of %d (the reference "
                           "stops the same way at its NDMAX)", b, 2 * S->nd[b], EB_SIM_NDMAX),
              EB_ERR_UNSUPPORTED);
    }
  if (icut) {
    // DOUBLE every flagged body, rebuild the bodies and F for the new lengths (electric.f:179-199)
    std::vector<int> nd_new(S->nd), off_new(S->k + 1, 0);
    for (int b = 0; b < S->k; ++b) {
      if (dbl[b]) nd_new[b] *= 2;
      off_new[b + 1] = off_new[b] + nd_new[b];
    }
    double2* ynew = nullptr;
    int *d_dbl = nullptr, *d_off = nullptr;
    if ((rc = dalloc(&ynew, (size_t)off_new[S->k])) || (rc = dalloc(&d_dbl, S->k)) || (rc = dalloc(&d_off, S->k + 1)))
      return rc;
    if (!upload(d_dbl, dbl, st) || !upload(d_off, off_new, st)) return (eb_set_error("eb_sim_step: upload failed"), EB_ERR_HIP);
    hipLaunchKernelGGL(k_sim_double, dim3(S->k), dim3(SP_THREADS), lds_bytes(B, 2), st, view(B),
                       (const double2*)S->ybar, d_dbl, d_off, ynew);
    EB_HIP_CHECK(hipGetLastError());
    EB_HIP_CHECK(hipStreamSynchronize(st));
    (void)hipFree(d_dbl);
    (void)hipFree(d_off);
    sim_free_state(S);
    S->nd = nd_new;
    if ((rc = sim_alloc(S)) != EB_OK) {
      (void)hipFree(ynew);
      return rc;
    }
    EB_HIP_CHECK(hipMemcpyAsync(S->ybar, ynew, sizeof(double2) * S->nbk, hipMemcpyDeviceToDevice, st));
    EB_HIP_CHECK(hipStreamSynchronize(st));
    (void)hipFree(ynew);
    S->nplot *= 2;
    ++S->doublings;
  }
  // the step law (electric.f:219-244)
  double dtmin = 1e10;
  for (int b = 0; b < S->k; ++b) dtmin = std::min(dtmin, S->cmin * S->arcl[b] / S->nd[b]);
  S->dt = dtmin;
  S->erra = fabs(S->areatot0 - S->areatot) / S->areatot0;
  ++S->steps;
  if (done) *done = S->time < S->tfinal - 1e-8 ? 0 : 1;
  return EB_OK;
}

extern "C" int eb_sim_get_state(eb_sim S, eb_sim_state* st) {
  if (!S || !st) return (eb_set_error("eb_sim_get_state: bad arguments"), EB_ERR_ARG);
  st->time = S->time;
  st->dt = S->dt;
  st->area_total = S->areatot;
  st->area_total0 = S->areatot0;
  st->area_error = S->erra;
  st->k = S->k;
  st->nbk = S->nbk;
  st->steps = S->steps;
  st->doublings = S->doublings;
  st->nd = S->nd.data();
  st->area = S->area.data();
  st->arcl = S->arcl.data();
  st->arcl0 = S->arcl0.data();
  st->ybar = S->ybar;
  return EB_OK;
}

extern "C" int eb_sim_copy_state(eb_sim S, double* dst) {
  if (!S || !dst) return (eb_set_error("eb_sim_copy_state: bad arguments"), EB_ERR_ARG);
  EB_HIP_CHECK(hipMemcpyAsync(dst, S->ybar, sizeof(double) * 2 * (size_t)S->nbk, hipMemcpyDefault, S->st));
  EB_HIP_CHECK(hipStreamSynchronize(S->st));
  return EB_OK;
}

extern "C" int eb_sim_dump(eb_sim S, const char* path) {
  if (!S || !path) return (eb_set_error("eb_sim_dump: bad arguments"), EB_ERR_ARG);
  std::vector<double> y(2 * (size_t)S->nbk);
  EB_HIP_CHECK(hipMemcpyAsync(y.data(), S->ybar, sizeof(double) * y.size(), hipMemcpyDeviceToHost, S->st));
  EB_HIP_CHECK(hipStreamSynchronize(S->st));
  FILE* f = fopen(path, "w");
  if (!f) return (eb_set_error("eb_sim_dump: cannot write %s", path), EB_ERR_ARG);
  // DUMPDATA (electric.f:1360-1388): list-directed integers, then (e22.16,2x,e22.16) records
  fprintf(f, " %d %d %d\n", S->k, S->nbk, 2 * S->nbk);
  for (int b = 0; b < S->k; ++b) fprintf(f, " %d", S->nd[b]);
  fprintf(f, "\n");
  // (the values carry 17 significant digits, one more than E22.16; READDATA reads both)
  auto rec = [&](double a, double b) { fprintf(f, "%22.16E  %22.16E\n", a, b); };
  rec(S->dt, S->time);
  for (int b = 0; b < S->k; ++b) fprintf(f, "%22.16E\n", S->area0[b]);
  for (int b = 0; b < S->k; ++b) fprintf(f, "%22.16E\n", S->arcl0[b]);
  for (size_t i = 0; i < (size_t)S->nbk; ++i) rec(y[2 * i], y[2 * i + 1]);
  fclose(f);
  return EB_OK;
}

// Fortran list-directed reading: values separated by blanks or commas, 'd'/'D' exponents;
// each READ starts on a new record and skips the rest of its last record
namespace {
struct ListReader {
  FILE* f;
  bool next_record(std::vector<std::string>& tok) {
    char line[4096];
    if (!fgets(line, sizeof line, f)) return false;
    tok.clear();
    std::string s(line);
    for (char& ch : s)
      if (ch == ',' || ch == '\t' || ch == '\r' || ch == '\n') ch = ' ';
    std::istringstream is(s);
    std::string t;
    while (is >> t) {
      for (char& ch : t)
        if (ch == 'd' || ch == 'D') ch = 'e';
      tok.push_back(t);
    }
    return true;
  }
  bool read(std::vector<double>& v, size_t count) {
    v.clear();
    std::vector<std::string> tok;
    while (v.size() < count) {
      if (!next_record(tok)) return false;
      for (const std::string& t : tok) {
        if (v.size() >= count) break;
        char* end = nullptr;
        const double x = strtod(t.c_str(), &end);
        if (end == t.c_str()) return false;
        v.push_back(x);
      }
    }
    return true;
  }
};
}  // namespace

extern "C" int eb_sim_read_bubbin(const char* path, eb_sim_params* p) {
  if (!path || !p) return (eb_set_error("eb_sim_read_bubbin: bad arguments"), EB_ERR_ARG);
  FILE* f = fopen(path, "r");
  if (!f) return (eb_set_error("eb_sim_read_bubbin: cannot open %s", path), EB_ERR_ARG);
  ListReader r{f};
  std::vector<double> v;
  memset(p, 0, sizeof(*p));
  bool ok = r.read(v, 1);
  if (ok) p->k = (int)v[0];
  ok = ok && p->k >= 1 && p->k <= EB_SIM_KMAX && r.read(v, (size_t)p->k);
  if (ok)
    for (int b = 0; b < p->k; ++b) p->nd[b] = (int)v[b];
  ok = ok && r.read(v, 3);
  if (ok) {
    p->dt = v[0];
    p->tfinal = v[1];
    p->nplot = (int)v[2];
  }
  ok = ok && r.read(v, 4);
  if (ok) {
    p->cap = v[0];
    p->bondx = v[1];
    p->bondy = v[2];
    p->e0 = v[3];
  }
  for (int b = 0; ok && b < p->k; ++b) {
    ok = r.read(v, 6);
    if (ok) {
      p->ai[b] = v[0];
      p->bi[b] = v[1];
      p->thetax[b] = v[2];
      p->cx[b] = v[3];
      p->cy[b] = v[4];
      p->ncyc[b] = (int)v[5];
    }
  }
  fclose(f);
  if (!ok) return (eb_set_error("eb_sim_read_bubbin: %s is not a complete run description", path), EB_ERR_ARG);
  return EB_OK;
}

extern "C" int eb_sim_restore(eb_sim* out, const eb_sim_params* p, const char* path, void* stream) {
  if (!out || !p || !path) return (eb_set_error("eb_sim_restore: bad arguments"), EB_ERR_ARG);
  *out = nullptr;
  FILE* f = fopen(path, "r");
  if (!f) return (eb_set_error("eb_sim_restore: cannot open %s", path), EB_ERR_ARG);
  ListReader r{f};
  std::vector<double> v;
  eb_sim_params q = *p;
  bool ok = r.read(v, 3);
  int nbk = 0;
  if (ok) {
    q.k = (int)v[0];
    nbk = (int)v[1];
  }
  ok = ok && q.k >= 1 && q.k <= EB_SIM_KMAX && r.read(v, (size_t)q.k);
  int sum = 0;
  if (ok)
    for (int b = 0; b < q.k; ++b) sum += (q.nd[b] = (int)v[b]);
  ok = ok && sum == nbk;
  double tstart = 0.0;
  ok = ok && r.read(v, 2);
  if (ok) {
    q.dt = v[0];
    tstart = v[1];
  }
  std::vector<double> area0(q.k), arcl0(q.k), z(2 * (size_t)nbk);
  for (int b = 0; ok && b < q.k; ++b)
    if ((ok = r.read(v, 1))) area0[b] = v[0];
  for (int b = 0; ok && b < q.k; ++b)
    if ((ok = r.read(v, 1))) arcl0[b] = v[0];
  for (int i = 0; ok && i < nbk; ++i)
    if ((ok = r.read(v, 2))) {
      z[2 * i] = v[0];
      z[2 * i + 1] = v[1];
    }
  fclose(f);
  if (!ok) return (eb_set_error("eb_sim_restore: %s is not a geometry.data file", path), EB_ERR_ARG);
  q.bondx = q.bondy = 0.0;
  int rc = check_params(&q);
  if (rc != EB_OK) return rc;
  eb_sim S = new eb_sim_s();
  S->p = q;
  S->st = (hipStream_t)stream;
  S->k = q.k;
  S->nd.assign(q.nd, q.nd + q.k);
  S->dt = q.dt;
  S->time = tstart;
  S->tfinal = tstart + p->tfinal;  // electric.f:88
  S->nplot = q.nplot;
  if ((rc = sim_alloc(S)) != EB_OK) {
    sim_free(S);
    return rc;
  }
  if (hipMemcpyAsync(S->ybar, z.data(), sizeof(double) * z.size(), hipMemcpyHostToDevice, S->st) != hipSuccess ||
      hipStreamSynchronize(S->st) != hipSuccess) {
    sim_free(S);
    return (eb_set_error("eb_sim_restore: upload failed"), EB_ERR_HIP);
  }
  S->area0 = area0;
  S->arcl0 = arcl0;
  S->area = area0;
  S->arcl = arcl0;
  sim_cmin(S);
  *out = S;
  return EB_OK;
}

extern "C" int eb_sim_destroy(eb_sim S) {
  sim_free(S);
  return EB_OK;
}
