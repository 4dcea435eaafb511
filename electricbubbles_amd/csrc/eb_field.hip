// eb_field.hip -- F on the device (include/eb_field.h): the Laplace solve, the
// Dirichlet-to-Neumann map, SETVELF and INTERFACENEW around the Stokes solve.
//
// Every layer potential runs on the GPU FMM (eb_fmm_plan): the Laplace operator's Hilbert
// product (DAPIF2, IFLAG7 = 3) and PVINTEV's on the Hilbert variant, INTERFACENEW's CADAPMOD on
// the MOD variant, SOLVEI's CADAPDR inside eb_matvec.  Per-body spectral work (interpolation
// to twice the points, FDIFFF, FOURINT) runs one workgroup per body with the in-LDS FFT.
#include <math.h>

#include <algorithm>
#include <vector>

#include "../../include/eb_field.h"
#include "../../include/eb_gmres.h"
#include "eb_bodies.h"

using namespace eb;

namespace {

constexpr double PI = 3.14159265358979323846;
constexpr int EW = 256;  // threads of the pointwise kernels

inline int ew_blocks(int n) { return (n + EW - 1) / EW; }

// NTERMS = -NINT(ln TOL / ln 2) (cadapdrop.f:126)
inline int nterms_for(double tol) { return -(int)lround(log(tol) / log(2.0)); }

// ------------------------------------------------------------------------------------------
// Laplace operator FASMVP_LAP (laplace.f:159-262)
// ------------------------------------------------------------------------------------------

__global__ void k_real_dens(int n, const double* __restrict__ u, double2* __restrict__ d) {
  const int i = blockIdx.x * EW + threadIdx.x;
  if (i < n) d[i] = make_double2(u[i], 0.0);
}

// w = u + 2 (self u - Re(CF / (2 pi i))) - 2 h u dsdth, self = h kappa dsdth / (4 pi)
// (laplace.f:225-237); CF = the Hilbert product of the charges u dz h
__global__ void k_lap_point(int n, const int* __restrict__ body_of, const double* __restrict__ h,
                            const double* __restrict__ u, const double2* __restrict__ cf,
                            const double* __restrict__ rkappa, const double* __restrict__ dsdth,
                            double* __restrict__ w) {
  const int i = blockIdx.x * EW + threadIdx.x;
  if (i >= n) return;
  const double hs = h[body_of[i]];
  const double self = 0.25 * hs * rkappa[i] * dsdth[i] / PI;
  // Re(z2pii * cf) with z2pii = 1 / (2 pi i) = -i / (2 pi)
  const double zc = self * u[i] - cf[i].y / (2.0 * PI);
  const double wi = u[i] + 2.0 * zc;
  const double du = hs * u[i] * dsdth[i];
  w[i] = wi - 2.0 * du;
}

// GETBCs (laplace.f:14-18): rhs = -2 Re z
__global__ void k_getbcs(int n, const double* __restrict__ x, double* __restrict__ rhs, double* __restrict__ soln) {
  const int i = blockIdx.x * EW + threadIdx.x;
  if (i >= n) return;
  const double r = -2.0 * x[i];
  rhs[i] = r;
  soln[i] = r;  // SOLVE_LAPL's initial guess soln = rhs (laplace.f:375-378)
}

// ------------------------------------------------------------------------------------------
// PVINTEV (pv_integral.f:53-212) and DIRICH_NEU_MAP (laplace.f:266-332)
// ------------------------------------------------------------------------------------------

// The spectrum of a length-N sequence in a (after the forward transform) becomes FINTERC's
// interpolant on 2N points (pv_integral.f:268-312): coefficients scaled by 1/N, the one at
// N/2 set to zero, the rest zero-padded.  Its values at the even fine points are the inverse
// length-N transform of those coefficients; at the odd fine points (the midpoints) the same
// with coefficient k' multiplied by exp(i pi k' / N), k' the signed wavenumber.
__device__ __forceinline__ void interp_coeffs(double2* a, double2* c, int N) {
  const double invn = 1.0 / N;
  const int nyk = N / 2;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    double2 v = make_double2(a[i].x * invn, a[i].y * invn);
    if (i == nyk) v = make_double2(0.0, 0.0);
    a[i] = v;
    const int ks = (i <= nyk) ? i : i - N;  // signed wavenumber
    double sn, cs;
    sincospi((double)ks / (double)N, &sn, &cs);
    c[i] = make_double2(v.x * cs - v.y * sn, v.x * sn + v.y * cs);
  }
  __syncthreads();
}

// The same for the curve z = x + i y, with the interpolant that PVINTEV forms for x and for y
// with FINTER, the real-data routine (pv_integral.f:216-252).  FINTER clears the imaginary part
// of coefficient N/2 - 1 and the real part of coefficient N/2 (the two words `work(ipt+2*nyk-1)`
// and `work(ipt+2*nyk)`, pv_integral.f:236-237) where FINTERC clears the Nyquist entry, and
// keeps the real part of the inverse transform.  For real data that halves the imaginary part of
// the +-(N/2 - 1) pair of x's interpolant (and of y's); the Nyquist entry is zero up to rounding.
// In terms of zhat = FFT(x + i y) only entries N/2 - 1 and N/2 + 1 change.  Resolved curves
// carry nothing at mode N/2 - 1; curves that need DOUBLE do.
__device__ __forceinline__ void interp_coeffs_curve(double2* a, double2* c, int N) {
  const double invn = 1.0 / N;
  const int nyk = N / 2;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const double2 v = make_double2(a[i].x * invn, a[i].y * invn);
    a[i] = i == nyk ? make_double2(0.0, 0.0) : v;
  }
  __syncthreads();
  if (threadIdx.x == 0 && nyk >= 2) {
    const double2 zp = a[nyk - 1], zm = a[nyk + 1];  // modes +(N/2 - 1) and -(N/2 - 1)
    // xhat = (zp + conj(zm)) / 2, yhat = (zp - conj(zm)) / (2 i), at mode +(N/2 - 1)
    const double2 xh = make_double2(0.5 * (zp.x + zm.x), 0.5 * (zp.y - zm.y));
    const double2 yh = make_double2(0.5 * (zp.y + zm.y), -0.5 * (zp.x - zm.x));
    // FINTER's interpolant: the imaginary parts halved; then zhat = xhat + i yhat at +(N/2 - 1)
    // and conj(xhat) + i conj(yhat) at -(N/2 - 1)
    const double2 xq = make_double2(xh.x, 0.5 * xh.y), yq = make_double2(yh.x, 0.5 * yh.y);
    a[nyk - 1] = make_double2(xq.x - yq.y, xq.y + yq.x);
    a[nyk + 1] = make_double2(xq.x + yq.y, -xq.y + yq.x);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const double2 v = a[i];
    const int ks = (i <= nyk) ? i : i - N;  // signed wavenumber
    double sn, cs;
    sincospi((double)ks / (double)N, &sn, &cs);
    c[i] = make_double2(v.x * cs - v.y * sn, v.x * sn + v.y * cs);
  }
  __syncthreads();
}

// Per body: the fine coordinates x2, y2 (2N points: even = the interpolated original points,
// odd = the midpoints) and the fine density: ucmplx = u dz interpolated at the midpoints, 0 at
// the original points (QA of pv_integral.f:160-163 is U2 h there, with h the FMM weight).
__global__ __launch_bounds__(SP_THREADS) void k_pv_interp(BodyView B, const double* __restrict__ x,
                                                           const double* __restrict__ y,
                                                           const double* __restrict__ ulap,
                                                           const double2* __restrict__ dz, double* __restrict__ x2,
                                                           double* __restrict__ y2, double2* __restrict__ uf) {
  extern __shared__ double2 lds[];
  const BodyCtx bc = body_ctx(B);
  const int N = bc.N, o0 = bc.o0, f0 = 2 * bc.o0;
  double2* a = lds;
  double2* c = lds + B.nmax;
  double2* ws = lds + 2 * B.nmax;
  // the curve, as one complex signal x + i y (FINTER of x and of y, pv_integral.f:156-159)
  for (int i = threadIdx.x; i < N; i += blockDim.x) a[i] = make_double2(x[o0 + i], y[o0 + i]);
  __syncthreads();
  fft(a, ws, bc, B, -1);
  interp_coeffs_curve(a, c, N);
  fft(a, ws, bc, B, +1);
  fft(c, ws, bc, B, +1);
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    x2[f0 + 2 * i] = a[i].x;
    y2[f0 + 2 * i] = a[i].y;
    x2[f0 + 2 * i + 1] = c[i].x;
    y2[f0 + 2 * i + 1] = c[i].y;
  }
  __syncthreads();
  // the density u dz (FINTERC, pv_integral.f:154-155), needed at the midpoints only
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const double2 d = dz[o0 + i];
    const double u = ulap[o0 + i];
    a[i] = make_double2(u * d.x, u * d.y);
  }
  __syncthreads();
  fft(a, ws, bc, B, -1);
  interp_coeffs(a, c, N);
  fft(c, ws, bc, B, +1);
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    uf[f0 + 2 * i] = make_double2(0.0, 0.0);
    uf[f0 + 2 * i + 1] = c[i];
  }
}

// fine weights: H(body) at the midpoints, 0 at the original points
__global__ void k_pv_weights(BodyView B, double2* __restrict__ wf) {
  const int k = blockIdx.x;
  const int N = B.nd[k], f0 = 2 * B.off[k];
  const double h = 2.0 * PI / N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    wf[f0 + 2 * i] = make_double2(0.0, 0.0);
    wf[f0 + 2 * i + 1] = make_double2(h, 0.0);
  }
}

// DIRICH_NEU_MAP (laplace.f:297-325): PHICMPLX = -z2pii CFIELD at the original points,
// zphi = -u/2 - PHICMPLX - z, phi = Re, psi = Im, dphi_dn = -Re(FDIFFF(psi)) / dsdth
__global__ __launch_bounds__(SP_THREADS) void k_dn_post(BodyView B, const double* __restrict__ x,
                                                         const double* __restrict__ y,
                                                         const double* __restrict__ ulap,
                                                         const double2* __restrict__ cf2,
                                                         const double* __restrict__ dsdth, double* __restrict__ phi,
                                                         double* __restrict__ psi, double* __restrict__ dphi) {
  extern __shared__ double2 lds[];
  const BodyCtx bc = body_ctx(B);
  const int N = bc.N, o0 = bc.o0, f0 = 2 * bc.o0;
  double2* a = lds;
  double2* ws = lds + B.nmax;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const double2 cf = cf2[f0 + 2 * i];
    // PHICMPLX = -z2pii CF = (i / (2 pi)) CF
    const double2 pc = make_double2(-cf.y / (2.0 * PI), cf.x / (2.0 * PI));
    const double zr = -0.5 * ulap[o0 + i] - pc.x - x[o0 + i];
    const double zi = -pc.y - y[o0 + i];
    if (phi) phi[o0 + i] = zr;
    if (psi) psi[o0 + i] = zi;
    a[i] = make_double2(zi, 0.0);
  }
  __syncthreads();
  fft(a, ws, bc, B, -1);
  const double invn = 1.0 / N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const double w = fdiff_k(i, N) * invn;
    const double2 v = a[i];
    a[i] = make_double2(-w * v.y, w * v.x);
  }
  __syncthreads();
  fft(a, ws, bc, B, +1);
  for (int i = threadIdx.x; i < N; i += blockDim.x) dphi[o0 + i] = -a[i].x / dsdth[o0 + i];
}

// ------------------------------------------------------------------------------------------
// SETVELF (stokesbubb.f:4-135) with zbond = 0 and iwall = 0: the bond-number and C_k terms
// are exactly zero then (zfi = 0, zc = 0), which leaves
//   rhs = -dz / (2 dsda) - zstr + (i E0 / 4) (alpha zf1 + zf(i) - zf(1)),
// zstr = -conj(zpsi), zpsi = -i cap z / 2 (stokesbubb.f:63-67), zf = the FOURINT integral of
// FFT(dz dphi_dn^2) / N, and alpha the value the C_k loop left behind (stokesbubb.f:94):
// 2 pi (N_k - 1) / N_k of the last body.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(SP_THREADS) void k_setvelf(BodyView B, const double* __restrict__ x,
                                                         const double* __restrict__ y, const double2* __restrict__ dz,
                                                         const double* __restrict__ dsda,
                                                         const double* __restrict__ dphi,
                                                         const double* __restrict__ cap_b,
                                                         const double* __restrict__ e0_b,
                                                         const double* __restrict__ alpha_b, double2* __restrict__ rhs) {
  extern __shared__ double2 lds[];
  const BodyCtx bc = body_ctx(B);
  const double cap = cap_b[bc.k], e0 = e0_b[bc.k], alpha_last = alpha_b[bc.k];
  const int N = bc.N, o0 = bc.o0, n = N / 2;
  double2* a = lds;
  double2* c = lds + B.nmax;
  double2* ws = lds + 2 * B.nmax;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const double d2 = dphi[o0 + i] * dphi[o0 + i];
    const double2 d = dz[o0 + i];
    a[i] = make_double2(d.x * d2, d.y * d2);
  }
  __syncthreads();
  fft(a, ws, bc, B, -1);
  const double invn = 1.0 / N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const double2 v = make_double2(a[i].x * invn, a[i].y * invn);
    a[i] = v;
    // FOURINT (electric.f:1307-1331)
    double2 r = make_double2(0, 0);
    if (i >= 1 && i <= n - 1) {
      const double ii = (double)i;
      r = make_double2(v.y / ii, -v.x / ii);
    } else if (i >= N - n + 1 && i <= N - 1) {
      const double ii = (double)(N - i);
      r = make_double2(-v.y / ii, v.x / ii);
    }
    c[i] = r;
  }
  __syncthreads();
  fft(c, ws, bc, B, +1);
  const double2 zf1 = a[0], c0 = c[0];
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const int g = o0 + i;
    const double2 d = dz[g];
    const double s = dsda[g];
    // zn = -0.5 dz / dsda; rhs = zn + zgrav (= 0) - zstr, zstr = -conj(zpsi)
    const double2 zn = make_double2(-0.5 * d.x / s, -0.5 * d.y / s);
    const double hc = 0.5 * cap;
    double rx = zn.x + hc * y[g];
    double ry = zn.y + hc * x[g];
    // electric term: (i E0 / 4) (alpha zf1 + zf(i) - zf(1))
    const double2 zs = make_double2(alpha_last * zf1.x + c[i].x - c0.x, alpha_last * zf1.y + c[i].y - c0.y);
    const double q = 0.25 * e0;
    rx += -q * zs.y;
    ry += q * zs.x;
    rhs[g] = make_double2(rx, ry);
  }
}

// ------------------------------------------------------------------------------------------
// INTERFACENEW (electric.f:851-878) with zc = 0 and iwall = 0
// ------------------------------------------------------------------------------------------
__global__ void k_interface_point(int n, const int* __restrict__ body_of, const double* __restrict__ h,
                                  const double* __restrict__ x, const double* __restrict__ y,
                                  const double2* __restrict__ dz, const double* __restrict__ dsda,
                                  const double* __restrict__ rkappa, const double2* __restrict__ ckern,
                                  const double2* __restrict__ u, const double2* __restrict__ udot,
                                  const double2* __restrict__ gw, const double* __restrict__ cap_b,
                                  double* __restrict__ un) {
  const int i = blockIdx.x * EW + threadIdx.x;
  if (i >= n) return;
  const int kb = body_of[i];
  const double hk = h[kb], cap = cap_b[kb];
  const double2 d = dz[i];
  const double s = dsda[i];
  const double2 zn = make_double2(d.y / s, -d.x / s);  // -i dz / dsda
  const double2 ui = u[i], ud = udot[i];
  // zcorr = (h Re(ckern) u - 2 h udot) / (2 pi i)
  const double r1 = 0.5 * hk * 2.0 * ckern[i].x;
  const double2 t = make_double2(r1 * ui.x - 2.0 * hk * ud.x, r1 * ui.y - 2.0 * hk * ud.y);
  const double2 zcorr = make_double2(t.y / (2.0 * PI), -t.x / (2.0 * PI));
  // zself = conj(u) h |dz| kappa zn^2 / (2 pi)
  const double f = hk * hypot(d.x, d.y) * 0.5 * rkappa[i];
  const double2 cu = make_double2(ui.x * f, -ui.y * f);
  const double2 zn2 = cmul(zn, zn);
  const double2 zs = cmul(cu, zn2);
  const double2 zself = make_double2(zs.x / PI, zs.y / PI);
  double2 g = gw[i];
  g = make_double2(g.x + zself.x + zcorr.x, g.y + zself.y + zcorr.y);
  // gradwr = conj(zpsi), zpsi = -i cap z / 2
  const double hc = 0.5 * cap;
  g = make_double2(g.x + hc * y[i], g.y + hc * x[i]);
  // zvel = -i gradw; Un = Re(zvel conj(zn))
  const double2 zv = make_double2(g.y, -g.x);
  un[i] = zv.x * zn.x + zv.y * zn.y;
}

__global__ void k_body_of_f(const int* __restrict__ off, int nbody, int* __restrict__ body_of) {
  const int k = blockIdx.x;
  for (int i = off[k] + threadIdx.x; i < off[k + 1]; i += blockDim.x) body_of[i] = k;
}

__global__ void k_wdz_f(int n, const double2* __restrict__ dz, const double* __restrict__ h,
                        const int* __restrict__ body_of, double2* __restrict__ wdz) {
  const int i = blockIdx.x * EW + threadIdx.x;
  if (i >= n) return;
  const double hk = h[body_of[i]];
  wdz[i] = make_double2(dz[i].x * hk, dz[i].y * hk);
}

}  // namespace

struct eb_field_s {
  eb_bodies B = nullptr;
  int k = 0, n = 0;
  double cap = 0.0, e0 = 0.0, alpha_last = 0.0;
  hipStream_t st = nullptr;
  bool have_geom = false;
  int mode = EB_MODE_AUTO;  // evaluation mode of every layer potential (eb_field_set_mode)
  int32_t it_lap = 0, ie_lap = 0, it_st = 0, ie_st = 0;  // the last call's GMRES counts
  int nt_lap = 37, nt_pv = 33, nt_mod = 43;  // TOL 1e-11 (laplace.f:212), 1e-10 (pv_integral.f:175), 1e-13
  // geometry
  double *x = nullptr, *y = nullptr, *tha = nullptr, *dsda = nullptr, *rkappa = nullptr, *area = nullptr,
         *xls = nullptr, *h = nullptr, *cap_b = nullptr, *e0_b = nullptr, *alpha_b = nullptr;
  double2 *dz = nullptr, *zk = nullptr, *ckern = nullptr, *wdz = nullptr;
  int* body_of = nullptr;
  // Laplace
  eb_fmm_plan lap = nullptr, pv = nullptr;
  eb_gmres glap = nullptr;
  double2 *dens = nullptr, *cf = nullptr, *uf = nullptr, *wf = nullptr, *cf2 = nullptr;
  double *x2 = nullptr, *y2 = nullptr, *rhs_lap = nullptr, *ulap = nullptr, *phi = nullptr, *psi = nullptr,
         *dphi = nullptr;
  // Stokes and the interface velocity
  eb_matvec mv = nullptr;
  eb_gmres gst = nullptr;
  eb_fmm_plan itf = nullptr;
  double *rhs_st = nullptr, *pk = nullptr, *un = nullptr;
  double2 *u = nullptr, *za = nullptr, *udot = nullptr, *gw = nullptr;
  // the captured RK2 step (eb_field_rk2_graph): device-controlled solves of the one system
  eb_gmres_batch glap1 = nullptr, gst1 = nullptr;
  hipGraphExec_t gexec = nullptr;
  double* g_ybar = nullptr;
  double g_dt = 0.0;
  int g_budget = 0, g_order = 0;
  double *ysave = nullptr, *yst = nullptr, *phi0 = nullptr, *dy = nullptr;
};

static void drop_graph(eb_field F) {
  if (F->gexec) (void)hipGraphExecDestroy(F->gexec);
  F->gexec = nullptr;
}

static void free_field(eb_field F) {
  if (!F) return;
  if (F->st) (void)hipStreamSynchronize(F->st);
  drop_graph(F);
  for (double* q : {F->ysave, F->yst, F->phi0, F->dy})
    if (q) (void)hipFree(q);
  if (F->glap1) eb_gmres_batch_destroy(F->glap1);
  if (F->gst1) eb_gmres_batch_destroy(F->gst1);
  if (F->lap) eb_fmm_plan_destroy(F->lap);
  if (F->pv) eb_fmm_plan_destroy(F->pv);
  if (F->itf) eb_fmm_plan_destroy(F->itf);
  if (F->mv) eb_matvec_destroy(F->mv);
  if (F->glap) eb_gmres_destroy(F->glap);
  if (F->gst) eb_gmres_destroy(F->gst);
  void* p[] = {F->cap_b, F->e0_b, F->alpha_b,
               F->x,    F->y,   F->tha,  F->dsda, F->rkappa, F->area,   F->xls,  F->h,   F->dz,   F->zk,
               F->ckern, F->wdz, F->body_of, F->dens, F->cf,  F->uf,     F->wf,   F->cf2, F->x2,   F->y2,
               F->rhs_lap, F->ulap, F->phi, F->psi, F->dphi, F->rhs_st, F->pk, F->un, F->u, F->za, F->udot, F->gw};
  for (void* q : p)
    if (q) (void)hipFree(q);
  delete F;
}

extern "C" int eb_field_create(eb_field* out, eb_bodies B, double cap, double e0, void* stream) {
  if (!out || !B) return (eb_set_error("eb_field_create: bad arguments"), EB_ERR_ARG);
  *out = nullptr;
  for (int k = 0; k < B->nbody; ++k)
    if (B->nd[k] > 2048)
      return (eb_set_error("eb_field_create: body %d has %d points (PVINTEV's interpolation takes up to 2048)", k,
                           B->nd[k]),
              EB_ERR_ARG);
  eb_field F = new eb_field_s();
  F->B = B;
  F->k = B->nbody;
  F->n = B->n;
  F->cap = cap;
  F->e0 = e0;
  F->st = (hipStream_t)stream;
  const int nl = B->nd[B->nbody - 1];
  F->alpha_last = 2.0 * PI * (nl - 1.0) / nl;  // stokesbubb.f:94, last body, last point
  F->nt_lap = nterms_for(1.0e-11);
  F->nt_pv = nterms_for(1.0e-10);
  F->nt_mod = nterms_for(1.0e-13);
  const int n = F->n, k = F->k;
  std::vector<double> h(k), capv(k, cap), e0v(k, e0), alv(k, F->alpha_last);
  for (int b = 0; b < k; ++b) h[b] = 2.0 * PI / B->nd[b];
  int rc;
  if ((rc = dalloc(&F->cap_b, k)) || (rc = dalloc(&F->e0_b, k)) || (rc = dalloc(&F->alpha_b, k)) ||
      (rc = dalloc(&F->x, n)) || (rc = dalloc(&F->y, n)) || (rc = dalloc(&F->tha, n)) || (rc = dalloc(&F->dsda, n)) ||
      (rc = dalloc(&F->rkappa, n)) || (rc = dalloc(&F->area, k)) || (rc = dalloc(&F->xls, 2)) ||
      (rc = dalloc(&F->h, k)) || (rc = dalloc(&F->dz, n)) || (rc = dalloc(&F->zk, k)) || (rc = dalloc(&F->ckern, n)) ||
      (rc = dalloc(&F->wdz, n)) || (rc = dalloc(&F->body_of, n)) || (rc = dalloc(&F->dens, n)) ||
      (rc = dalloc(&F->cf, n)) || (rc = dalloc(&F->uf, 2 * (size_t)n)) || (rc = dalloc(&F->wf, 2 * (size_t)n)) ||
      (rc = dalloc(&F->cf2, 2 * (size_t)n)) || (rc = dalloc(&F->x2, 2 * (size_t)n)) ||
      (rc = dalloc(&F->y2, 2 * (size_t)n)) || (rc = dalloc(&F->rhs_lap, n)) || (rc = dalloc(&F->ulap, n)) ||
      (rc = dalloc(&F->phi, n)) || (rc = dalloc(&F->psi, n)) || (rc = dalloc(&F->dphi, n)) ||
      (rc = dalloc(&F->rhs_st, 2 * (size_t)n)) || (rc = dalloc(&F->pk, k)) || (rc = dalloc(&F->un, n)) ||
      (rc = dalloc(&F->u, n)) || (rc = dalloc(&F->za, k)) || (rc = dalloc(&F->udot, n)) || (rc = dalloc(&F->gw, n))) {
    free_field(F);
    return rc;
  }
  if (!upload(F->h, h, F->st) || !upload(F->cap_b, capv, F->st) || !upload(F->e0_b, e0v, F->st) ||
      !upload(F->alpha_b, alv, F->st)) {
    free_field(F);
    return (eb_set_error("eb_field_create: upload failed"), EB_ERR_HIP);
  }
  hipLaunchKernelGGL(k_body_of_f, dim3(k), dim3(256), 0, F->st, B->d_off, k, F->body_of);
  hipLaunchKernelGGL(k_pv_weights, dim3(k), dim3(256), 0, F->st, view(B), F->wf);
  if ((rc = eb_gmres_create(&F->glap, n, 100, stream)) != EB_OK ||
      (rc = eb_gmres_create(&F->gst, 2 * n, 100, stream)) != EB_OK) {
    free_field(F);
    return rc;
  }
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(F->st) != hipSuccess) {
    free_field(F);
    return (eb_set_error("eb_field_create: kernel failed"), EB_ERR_HIP);
  }
  *out = F;
  return EB_OK;
}

extern "C" int eb_field_destroy(eb_field F) {
  free_field(F);
  return EB_OK;
}

// EB_MODE_AUTO's rule: the O(N^2) sum up to EB_DIRECT_MAX points, the FMM with LEVMAX 7 above;
// a mode set by eb_field_set_mode overrides it
static int plan_mode(eb_field F, int n) {
  if (F->mode != EB_MODE_AUTO) return F->mode;
  return n <= EB_DIRECT_MAX ? EB_MODE_DIRECT : EB_MODE_FMM;
}

extern "C" int eb_field_set_mode(eb_field F, int32_t mode) {
  if (!F || (mode != EB_MODE_AUTO && mode != EB_MODE_FMM && mode != EB_MODE_DIRECT))
    return (eb_set_error("eb_field_set_mode: bad arguments"), EB_ERR_ARG);
  if (mode == F->mode) return EB_OK;
  if (F->st) (void)hipStreamSynchronize(F->st);
  drop_graph(F);  // it holds the old plans
  // the plans are made again, in the new mode, by the next eb_field_set_geometry
  eb_fmm_plan* plans[] = {&F->lap, &F->pv, &F->itf};
  for (eb_fmm_plan* q : plans)
    if (*q) {
      eb_fmm_plan_destroy(*q);
      *q = nullptr;
    }
  if (F->mv) eb_matvec_destroy(F->mv);
  F->mv = nullptr;
  F->have_geom = false;
  F->mode = mode;
  return EB_OK;
}

extern "C" int eb_field_set_geometry(eb_field F, const double* rku) {
  if (!F || !rku) return (eb_set_error("eb_field_set_geometry: bad arguments"), EB_ERR_ARG);
  const int n = F->n;
  void* s = (void*)F->st;
  int rc = eb_extractdat(F->B, rku, F->x, F->y, (double*)F->zk, (double*)F->dz, F->tha, F->dsda, F->rkappa,
                         (double*)F->ckern, F->area, F->xls, s);
  if (rc != EB_OK) return rc;
  hipLaunchKernelGGL(k_wdz_f, dim3(ew_blocks(n)), dim3(EW), 0, F->st, n, F->dz, F->h, F->body_of, F->wdz);
  EB_HIP_CHECK(hipGetLastError());
  const double* wd = (const double*)F->wdz;
  if (!F->have_geom) {
    if ((rc = eb_fmm_plan_create(&F->lap, n, F->x, F->y, wd, 7, F->nt_lap, EB_VARIANT_HILBERT, plan_mode(F, n), s)) ||
        (rc = eb_fmm_plan_create(&F->itf, n, F->x, F->y, wd, 7, F->nt_mod, EB_VARIANT_MOD, plan_mode(F, n), s)) ||
        (rc = eb_matvec_create(&F->mv, F->B, F->x, F->y, (const double*)F->dz, F->dsda, F->rkappa,
                               (const double*)F->zk, 7, F->nt_mod, plan_mode(F, n), s)))
      return rc;
  } else {
    if ((rc = eb_fmm_plan_rebuild(F->lap, F->x, F->y, wd, s)) || (rc = eb_fmm_plan_rebuild(F->itf, F->x, F->y, wd, s)) ||
        (rc = eb_matvec_rebuild(F->mv, F->x, F->y, (const double*)F->dz, F->dsda, F->rkappa, (const double*)F->zk, s)))
      return rc;
  }
  F->have_geom = true;
  return EB_OK;
}

static int need_geom(eb_field F, const char* what) {
  if (!F || !F->have_geom) {
    eb_set_error("%s: no geometry (call eb_field_set_geometry or eb_field_rhs first)", what);
    return EB_ERR_ARG;
  }
  return EB_OK;
}

extern "C" int eb_field_laplace_apply(eb_field F, const double* u, double* w) {
  int rc = need_geom(F, "eb_field_laplace_apply");
  if (rc != EB_OK) return rc;
  const int n = F->n;
  hipLaunchKernelGGL(k_real_dens, dim3(ew_blocks(n)), dim3(EW), 0, F->st, n, u, F->dens);
  if ((rc = eb_fmm_apply(F->lap, (const double*)F->dens, (double*)F->cf, (void*)F->st)) != EB_OK) return rc;
  hipLaunchKernelGGL(k_lap_point, dim3(ew_blocks(n)), dim3(EW), 0, F->st, n, F->body_of, F->h, u, F->cf, F->rkappa,
                     F->dsda, w);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

static int lap_gmres_apply(void* ctx, const double* x, double* y, void* stream) {
  (void)stream;
  return eb_field_laplace_apply((eb_field)ctx, x, y);
}

// SOLVE_LAPL with the device-controlled DGMRES of one system (at most `budget` Arnoldi steps;
// no host synchronisation): the same steps and tests as eb_gmres_solve, the iteration count
// and IERR stay on the device (eb_gmres_batch_results).  The pair sums of the products after
// convergence are skipped through the plan's guard.
static int laplace_solve_async(eb_field F, double* ulap, int budget) {
  const int n = F->n;
  hipLaunchKernelGGL(k_getbcs, dim3(ew_blocks(n)), dim3(EW), 0, F->st, n, F->x, F->rhs_lap, ulap);
  EB_HIP_CHECK(hipGetLastError());
  eb_fmm_set_guard(F->lap, eb_gmres_batch_active(F->glap1));
  const int rc = eb_gmres_batch_solve_async(F->glap1, lap_gmres_apply, F, F->rhs_lap, ulap, 1.0e-10, budget);
  eb_fmm_set_guard(F->lap, nullptr);
  return rc;
}

extern "C" int eb_field_laplace_solve(eb_field F, double* ulap, int32_t* iter, int32_t* ierr) {
  int rc = need_geom(F, "eb_field_laplace_solve");
  if (rc != EB_OK || !ulap) return rc != EB_OK ? rc : (eb_set_error("eb_field_laplace_solve: bad arguments"), EB_ERR_ARG);
  const int n = F->n;
  hipLaunchKernelGGL(k_getbcs, dim3(ew_blocks(n)), dim3(EW), 0, F->st, n, F->x, F->rhs_lap, ulap);
  EB_HIP_CHECK(hipGetLastError());
  int32_t it = 0, ie = 0;
  double err = 0.0;
  // SOLVE_LAPL: ITOL 0, TOL 1e-10, IGWK(5) = -1, MAXL 100 (laplace.f:355-384; electric.f:1035)
  rc = eb_gmres_solve(F->glap, lap_gmres_apply, F, F->rhs_lap, ulap, 1.0e-10, -1, &it, &err, &ie);
  F->it_lap = it;
  F->ie_lap = ie;
  if (iter) *iter = it;
  if (ierr) *ierr = ie;
  return rc;
}

extern "C" int eb_field_dirich_neu(eb_field F, const double* ulap, double* phi, double* psi, double* dphi_dn) {
  int rc = need_geom(F, "eb_field_dirich_neu");
  if (rc != EB_OK) return rc;
  if (!ulap || !dphi_dn) return (eb_set_error("eb_field_dirich_neu: bad arguments"), EB_ERR_ARG);
  eb_bodies B = F->B;
  const int n2 = 2 * F->n;
  void* s = (void*)F->st;
  hipLaunchKernelGGL(k_pv_interp, dim3(B->nbody), dim3(SP_THREADS), lds_bytes(B, 2), F->st, view(B), F->x, F->y, ulap,
                     F->dz, F->x2, F->y2, F->uf);
  EB_HIP_CHECK(hipGetLastError());
  if (!F->pv)
    rc = eb_fmm_plan_create(&F->pv, n2, F->x2, F->y2, (const double*)F->wf, 7, F->nt_pv, EB_VARIANT_HILBERT,
                            plan_mode(F, n2), s);
  else
    rc = eb_fmm_plan_rebuild(F->pv, F->x2, F->y2, (const double*)F->wf, s);
  if (rc != EB_OK) return rc;
  if ((rc = eb_fmm_apply(F->pv, (const double*)F->uf, (double*)F->cf2, s)) != EB_OK) return rc;
  hipLaunchKernelGGL(k_dn_post, dim3(B->nbody), dim3(SP_THREADS), lds_bytes(B, 1), F->st, view(B), F->x, F->y, ulap,
                     F->cf2, F->dsda, phi, psi, dphi_dn);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" int eb_field_setvelf(eb_field F, const double* dphi_dn, double* rhs) {
  int rc = need_geom(F, "eb_field_setvelf");
  if (rc != EB_OK) return rc;
  if (!dphi_dn || !rhs) return (eb_set_error("eb_field_setvelf: bad arguments"), EB_ERR_ARG);
  eb_bodies B = F->B;
  hipLaunchKernelGGL(k_setvelf, dim3(B->nbody), dim3(SP_THREADS), lds_bytes(B, 2), F->st, view(B), F->x, F->y, F->dz,
                     F->dsda, dphi_dn, F->cap_b, F->e0_b, F->alpha_b, (double2*)rhs);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" int eb_field_stokes_solve(eb_field F, const double* rhs, double* u, double* za, double* pk, int32_t* iter,
                                     int32_t* ierr) {
  int rc = need_geom(F, "eb_field_stokes_solve");
  if (rc != EB_OK) return rc;
  int32_t it = 0, ie = 0;
  double err = 0.0;
  rc = eb_stokes_solve(F->mv, F->gst, rhs, u, za, pk, &it, &err, &ie, (void*)F->st);
  F->it_st = it;
  F->ie_st = ie;
  if (iter) *iter = it;
  if (ierr) *ierr = ie;
  return rc;
}

extern "C" int eb_field_interface(eb_field F, const double* u, double* un) {
  int rc = need_geom(F, "eb_field_interface");
  if (rc != EB_OK) return rc;
  if (!u || !un) return (eb_set_error("eb_field_interface: bad arguments"), EB_ERR_ARG);
  const int n = F->n;
  void* s = (void*)F->st;
  // udot = FDIFFF(u) per body (electric.f:820-825), GW = CADAPMOD (electric.f:835-837)
  if ((rc = eb_fdifff(F->B, u, (double*)F->udot, s)) != EB_OK) return rc;
  if ((rc = eb_fmm_apply(F->itf, u, (double*)F->gw, s)) != EB_OK) return rc;
  hipLaunchKernelGGL(k_interface_point, dim3(ew_blocks(n)), dim3(EW), 0, F->st, n, F->body_of, F->h, F->x, F->y,
                     F->dz, F->dsda, F->rkappa, F->ckern, (const double2*)u, F->udot, F->gw, F->cap_b, un);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

// SOLVEI with the device-controlled DGMRES of one system: soln = rhs as the initial guess,
// then za and pk of the solution (stokesbubb.f:343-409), as eb_stokes_solve
static int stokes_solve_async(eb_field F, int budget) {
  void* s = (void*)F->st;
  const size_t nu = 2 * (size_t)F->n;
  EB_HIP_CHECK(hipMemcpyAsync(F->u, F->rhs_st, sizeof(double) * nu, hipMemcpyDeviceToDevice, F->st));
  eb_matvec_set_guard(F->mv, eb_gmres_batch_active(F->gst1));
  int rc = eb_gmres_batch_solve_async(F->gst1, eb_matvec_gmres_apply, F->mv, F->rhs_st, (double*)F->u, 1.0e-10, budget);
  eb_matvec_set_guard(F->mv, nullptr);
  if (rc != EB_OK) return rc;
  return eb_matvec_parts(F->mv, (const double*)F->u, (double*)F->za, F->pk, s);
}

// F (electric.f:968-1109).  budget = 0: the host-controlled solves; budget > 0: the
// device-controlled ones with at most `budget` Arnoldi steps each and no host synchronisation
// (the graph-capturable form; the plans and the workspaces must exist already).
static int field_rhs(eb_field F, const double* rku, double* rkup, int budget) {
  int rc;
  int32_t it, ie;
  if ((rc = eb_field_set_geometry(F, rku)) != EB_OK) return rc;
  rc = budget > 0 ? laplace_solve_async(F, F->ulap, budget) : eb_field_laplace_solve(F, F->ulap, &it, &ie);
  if (rc != EB_OK) return rc;
  if ((rc = eb_field_dirich_neu(F, F->ulap, F->phi, F->psi, F->dphi)) != EB_OK) return rc;
  if ((rc = eb_field_setvelf(F, F->dphi, F->rhs_st)) != EB_OK) return rc;
  rc = budget > 0 ? stokes_solve_async(F, budget)
                  : eb_field_stokes_solve(F, F->rhs_st, (double*)F->u, (double*)F->za, F->pk, &it, &ie);
  if (rc != EB_OK) return rc;
  if ((rc = eb_field_interface(F, (const double*)F->u, F->un)) != EB_OK) return rc;
  return eb_ftail(F->B, F->tha, F->un, (const double*)F->dz, F->dsda, rkup, (void*)F->st);
}

extern "C" int eb_field_rhs(eb_field F, const double* rku, double* rkup) {
  if (!F || !rku || !rkup) return (eb_set_error("eb_field_rhs: bad arguments"), EB_ERR_ARG);
  return field_rhs(F, rku, rkup, 0);
}

// One Heun step (electric.f:773-783) on the field's buffers: phi0 = dt F(ybar), y = ybar + phi0,
// ybar += phi0/2 + dt F(y)/2, with the RK2 kernels' roundings.
static int field_rk2_step(eb_field F, double* ybar, double dt, int budget) {
  const int neq = 2 * F->n;
  void* s = (void*)F->st;
  int rc;
  if ((rc = field_rhs(F, ybar, F->dy, budget)) != EB_OK) return rc;
  if ((rc = eb_rk2_stage1(neq, dt, ybar, F->dy, F->phi0, F->yst, s)) != EB_OK) return rc;
  if ((rc = field_rhs(F, F->yst, F->dy, budget)) != EB_OK) return rc;
  return eb_rk2_stage2(neq, dt, F->phi0, F->dy, ybar, s);
}

// One classic RK4 step on the same buffers (phi0 holds the slopes' weighted sum, yst the stage
// states): four evaluations of F, each followed by its stage update.
static int field_rk4_step(eb_field F, double* ybar, double dt, int budget) {
  const int neq = 2 * F->n;
  void* s = (void*)F->st;
  int rc;
  for (int q = 0; q < 4; ++q) {
    if ((rc = field_rhs(F, q == 0 ? ybar : F->yst, F->dy, budget)) != EB_OK) return rc;
    if ((rc = eb_rk4_stage(neq, q, dt, ybar, F->dy, F->phi0, F->yst, s)) != EB_OK) return rc;
  }
  return EB_OK;
}

static int field_step(eb_field F, double* ybar, double dt, int budget, int order) {
  return order == 4 ? field_rk4_step(F, ybar, dt, budget) : field_rk2_step(F, ybar, dt, budget);
}

// eb_field_rk2_graph (order 2) and eb_field_rk4_graph (order 4)
static int field_step_graph(eb_field F, double* ybar, double dt, int32_t budget, int32_t* redone, int order,
                            const char* name) {
  if (!F || !ybar || budget < 1) return (eb_set_error("%s: bad arguments", name), EB_ERR_ARG);
  if (F->st == nullptr)
    return (eb_set_error("%s: the field must use a non-default stream (graph capture)", name), EB_ERR_ARG);
  const size_t neq = 2 * (size_t)F->n;
  int rc;
  if (!F->ysave) {
    if ((rc = dalloc(&F->ysave, neq)) || (rc = dalloc(&F->yst, neq)) || (rc = dalloc(&F->phi0, neq)) ||
        (rc = dalloc(&F->dy, neq)))
      return rc;
  }
  if (!F->glap1) {
    const int32_t ol[2] = {0, F->n}, ou[2] = {0, 2 * F->n};
    // MAXL 100 as SOLVE_LAPL and SOLVEI (laplace.f:355-384, electric.f:1011,1060)
    if ((rc = eb_gmres_batch_create(&F->glap1, 1, ol, 100, (void*)F->st)) != EB_OK ||
        (rc = eb_gmres_batch_create(&F->gst1, 1, ou, 100, (void*)F->st)) != EB_OK)
      return rc;
  }
  // the plans are made (with synchronisations) by an eager F before the first capture
  if ((!F->have_geom || !F->pv) && (rc = field_rhs(F, ybar, F->dy, 0)) != EB_OK) return rc;
  if (!F->gexec || F->g_ybar != ybar || F->g_dt != dt || F->g_budget != budget || F->g_order != order) {
    drop_graph(F);
    EB_HIP_CHECK(hipStreamSynchronize(F->st));
    hipGraph_t graph = nullptr;
    EB_HIP_CHECK(hipStreamBeginCapture(F->st, hipStreamCaptureModeThreadLocal));
    int crc = hipMemcpyAsync(F->ysave, ybar, sizeof(double) * neq, hipMemcpyDeviceToDevice, F->st) == hipSuccess
                  ? EB_OK
                  : EB_ERR_HIP;
    if (crc == EB_OK) crc = eb_gmres_batch_clear_fail(F->glap1);
    if (crc == EB_OK) crc = eb_gmres_batch_clear_fail(F->gst1);
    if (crc == EB_OK) crc = field_step(F, ybar, dt, budget, order);
    const hipError_t e = hipStreamEndCapture(F->st, &graph);
    if (crc != EB_OK || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      if (crc == EB_OK) eb_set_error("%s: capture failed: %s", name, hipGetErrorString(e));
      return crc != EB_OK ? crc : EB_ERR_HIP;
    }
    const hipError_t ie = hipGraphInstantiate(&F->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
      F->gexec = nullptr;
      return (eb_set_error("%s: instantiate failed: %s", name, hipGetErrorString(ie)), EB_ERR_HIP);
    }
    F->g_ybar = ybar;
    F->g_dt = dt;
    F->g_budget = budget;
    F->g_order = order;
  }
  EB_HIP_CHECK(hipGraphLaunch(F->gexec, F->st));
  int32_t fl = 0, fs = 0;
  if ((rc = eb_gmres_batch_results(F->glap1, &F->it_lap, nullptr, &F->ie_lap, &fl)) != EB_OK ||
      (rc = eb_gmres_batch_results(F->gst1, &F->it_st, nullptr, &F->ie_st, &fs)) != EB_OK)
    return rc;
  if (redone) *redone = 0;
  if (fl + fs > 0) {
    // a solve needed more than `budget` Arnoldi steps: the step again from the saved state
    // with the host-controlled solves
    EB_HIP_CHECK(hipMemcpyAsync(ybar, F->ysave, sizeof(double) * neq, hipMemcpyDeviceToDevice, F->st));
    if ((rc = field_step(F, ybar, dt, 0, order)) != EB_OK) return rc;
    EB_HIP_CHECK(hipStreamSynchronize(F->st));
    if (redone) *redone = 1;
  }
  return EB_OK;
}

extern "C" int eb_field_rk2_graph(eb_field F, double* ybar, double dt, int32_t budget, int32_t* redone) {
  return field_step_graph(F, ybar, dt, budget, redone, 2, "eb_field_rk2_graph");
}

extern "C" int eb_field_rk4_graph(eb_field F, double* ybar, double dt, int32_t budget, int32_t* redone) {
  return field_step_graph(F, ybar, dt, budget, redone, 4, "eb_field_rk4_graph");
}

extern "C" int eb_field_iterations(eb_field F, int32_t* lap_iter, int32_t* lap_ierr, int32_t* stokes_iter,
                                   int32_t* stokes_ierr) {
  if (!F) return (eb_set_error("eb_field_iterations: bad arguments"), EB_ERR_ARG);
  if (lap_iter) *lap_iter = F->it_lap;
  if (lap_ierr) *lap_ierr = F->ie_lap;
  if (stokes_iter) *stokes_iter = F->it_st;
  if (stokes_ierr) *stokes_ierr = F->ie_st;
  return EB_OK;
}

extern "C" int eb_field_get_geometry(eb_field F, eb_field_geometry* g) {
  if (!F || !g) return (eb_set_error("eb_field_get_geometry: bad arguments"), EB_ERR_ARG);
  g->x = F->x;
  g->y = F->y;
  g->dz = (const double*)F->dz;
  g->tha = F->tha;
  g->dsda = F->dsda;
  g->rkappa = F->rkappa;
  g->ckern = (const double*)F->ckern;
  g->zk = (const double*)F->zk;
  return EB_OK;
}

// ==========================================================================================
// F of many independent simulations at once (eb_sweep_field)
// ==========================================================================================
//
// The same chain for every simulation of a sweep, each stage one launch for all of them:
// the Laplace and Stokes solves run in the batched GMRES (one system per simulation), the
// layer potentials sum over each simulation's own points (the O(N^2) sum that the plain path
// also takes up to 2048 points), the per-body spectral kernels already cover every body.
namespace {

constexpr int HB_T = 256;

// CF_t = sum_s q_s / (z_t - z_s), q = dens * w, over one simulation: targets t and sources s
// are points base + stride * i + toff / soff of the simulation's segment [o, o + n) (the
// coarse points, stride 1 with the self pair dropped, or the fine points, stride 2: targets
// at the original points, sources at the midpoints).  One 256-thread block per (simulation,
// 256 targets) tile; sources staged through LDS.
template <bool SELF>
__global__ __launch_bounds__(HB_T) void k_hilbert_seg(const int2* __restrict__ tiles, const int* __restrict__ soff,
                                                      int stride, int toff, int srcoff, const double* __restrict__ x,
                                                      const double* __restrict__ y, const double2* __restrict__ dens,
                                                      const double2* __restrict__ w, const int* __restrict__ mask,
                                                      double2* __restrict__ cf) {
  __shared__ double2 sz[HB_T], sq[HB_T];
  const int2 tl = tiles[blockIdx.x];
  if (mask && !mask[tl.x]) return;  // a system the solver no longer iterates: output unused
  const int o = soff[tl.x], n = soff[tl.x + 1] - o;
  const int base = stride * o;
  const int it = tl.y + (int)threadIdx.x;
  const bool act = it < n;
  const int ti = base + stride * (act ? it : 0) + toff;
  const double zx = x[ti], zy = y[ti];
  double ax = 0.0, ay = 0.0;
  for (int c0 = 0; c0 < n; c0 += HB_T) {
    __syncthreads();
    const int js = c0 + (int)threadIdx.x;
    if (js < n) {
      const int sj = base + stride * js + srcoff;
      sz[threadIdx.x] = make_double2(x[sj], y[sj]);
      sq[threadIdx.x] = cmul(dens[sj], w[sj]);
    }
    __syncthreads();
    const int m = min(HB_T, n - c0);
    if (act)
      for (int jj = 0; jj < m; ++jj) {
        // q / (z_t - z_s) = q conj(d) / |d|^2, d = z_t - z_s
        const double dx = zx - sz[jj].x, dy = zy - sz[jj].y;
        const bool self = SELF && (c0 + jj == it);
        const double r2 = self ? 1.0 : fma(dx, dx, dy * dy);
        const double ir = self ? 0.0 : rcp_nr(r2);
        const double2 q = sq[jj];
        ax = fma(fma(q.x, dx, q.y * dy), ir, ax);
        ay = fma(fma(q.y, dx, -q.x * dy), ir, ay);
      }
  }
  if (act) cf[base + stride * it + toff] = make_double2(ax, ay);
}

}  // namespace

struct eb_sweep_field_s {
  eb_bodies B = nullptr;
  int nsim = 0, k = 0, n = 0, ntile = 0;
  hipStream_t st = nullptr;
  bool have_geom = false;
  std::vector<int> soff_h, uoff_h, sim_off;
  std::vector<int32_t> it_lap, ie_lap, it_st, ie_st;
  int *soff = nullptr, *body_of = nullptr;
  int2* tiles = nullptr;
  double *h = nullptr, *cap_b = nullptr, *e0_b = nullptr, *alpha_b = nullptr;
  double *x = nullptr, *y = nullptr, *tha = nullptr, *dsda = nullptr, *rkappa = nullptr, *area = nullptr,
         *xls = nullptr;
  double2 *dz = nullptr, *zk = nullptr, *ckern = nullptr, *wdz = nullptr;
  double2 *dens = nullptr, *cf = nullptr, *uf = nullptr, *wf = nullptr, *cf2 = nullptr;
  double *x2 = nullptr, *y2 = nullptr, *rhs_lap = nullptr, *ulap = nullptr, *dphi = nullptr;
  eb_gmres_batch glap = nullptr, gst = nullptr;
  eb_sweep sw = nullptr;
  double *rhs_st = nullptr, *pk = nullptr, *un = nullptr;
  double2 *u = nullptr, *za = nullptr, *udot = nullptr, *gw = nullptr;
  // the captured RK2 step (eb_sweep_field_rk2_graph)
  hipGraphExec_t gexec = nullptr;
  double* g_ybar = nullptr;
  double g_dt = 0.0;
  int g_budget = 0, g_order = 0;
  double *ysave = nullptr, *yst = nullptr, *phi0 = nullptr, *dy = nullptr;
  // EB_MODE_FMM (or AUTO with a simulation above EB_DIRECT_MAX points): the Hilbert products of
  // the Laplace operator and of PVINTEV on segmented FMM plans (one segment per simulation), and
  // the sweep operator's layer potentials likewise (eb_sweep_set_mode)
  int mode = EB_MODE_AUTO, maxn = 0;
  eb_fmm_plan slap = nullptr, spv = nullptr;
  std::vector<int> soff2_h;
};

static bool sf_fmm(eb_sweep_field F) {
  return F->mode == EB_MODE_FMM || (F->mode == EB_MODE_AUTO && F->maxn > EB_DIRECT_MAX);
}

static void free_sf(eb_sweep_field F) {
  if (!F) return;
  if (F->st) (void)hipStreamSynchronize(F->st);
  if (F->gexec) (void)hipGraphExecDestroy(F->gexec);
  if (F->slap) eb_fmm_plan_destroy(F->slap);
  if (F->spv) eb_fmm_plan_destroy(F->spv);
  for (double* q : {F->ysave, F->yst, F->phi0, F->dy})
    if (q) (void)hipFree(q);
  if (F->glap) eb_gmres_batch_destroy(F->glap);
  if (F->gst) eb_gmres_batch_destroy(F->gst);
  if (F->sw) eb_sweep_destroy(F->sw);
  void* p[] = {F->soff, F->body_of, F->tiles, F->h, F->cap_b, F->e0_b, F->alpha_b, F->x, F->y, F->tha, F->dsda,
               F->rkappa, F->area, F->xls, F->dz, F->zk, F->ckern, F->wdz, F->dens, F->cf, F->uf, F->wf, F->cf2,
               F->x2, F->y2, F->rhs_lap, F->ulap, F->dphi, F->rhs_st, F->pk, F->un, F->u, F->za, F->udot, F->gw};
  for (void* q : p)
    if (q) (void)hipFree(q);
  delete F;
}

extern "C" int eb_sweep_field_create(eb_sweep_field* out, eb_bodies B, int32_t nsim, const int32_t* sim_off,
                                     const double* cap, const double* e0, void* stream) {
  if (!out || !B || nsim < 1 || !sim_off || !cap || !e0)
    return (eb_set_error("eb_sweep_field_create: bad arguments"), EB_ERR_ARG);
  *out = nullptr;
  if (sim_off[0] != 0 || sim_off[nsim] != B->nbody)
    return (eb_set_error("eb_sweep_field_create: simulation offsets must run from 0 to the body count"), EB_ERR_ARG);
  for (int s = 0; s < nsim; ++s)
    if (sim_off[s + 1] <= sim_off[s])
      return (eb_set_error("eb_sweep_field_create: simulation %d has no body", s), EB_ERR_ARG);
  for (int b = 0; b < B->nbody; ++b)
    if (B->nd[b] > 2048)
      return (eb_set_error("eb_sweep_field_create: body %d has %d points (up to 2048)", b, B->nd[b]), EB_ERR_ARG);
  eb_sweep_field F = new eb_sweep_field_s();
  F->B = B;
  F->nsim = nsim;
  F->k = B->nbody;
  F->n = B->n;
  F->st = (hipStream_t)stream;
  const int n = F->n, k = F->k;
  F->sim_off.assign(sim_off, sim_off + nsim + 1);
  F->soff_h.resize(nsim + 1);
  F->uoff_h.resize(nsim + 1);
  for (int s = 0; s <= nsim; ++s) {
    F->soff_h[s] = B->off[sim_off[s]];
    F->uoff_h[s] = 2 * F->soff_h[s];
  }
  F->soff2_h = F->uoff_h;  // PVINTEV's points: twice each simulation's
  for (int s = 0; s < nsim; ++s) F->maxn = std::max(F->maxn, F->soff_h[s + 1] - F->soff_h[s]);
  std::vector<double> h(k), capv(k), e0v(k), alv(k);
  for (int s = 0; s < nsim; ++s) {
    const int nl = B->nd[sim_off[s + 1] - 1];
    for (int b = sim_off[s]; b < sim_off[s + 1]; ++b) {
      h[b] = 2.0 * PI / B->nd[b];
      capv[b] = cap[s];
      e0v[b] = e0[s];
      alv[b] = 2.0 * PI * (nl - 1.0) / nl;  // stokesbubb.f:94, the simulation's last body
    }
  }
  std::vector<int> ord(nsim);
  for (int s = 0; s < nsim; ++s) ord[s] = s;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
    return F->soff_h[a + 1] - F->soff_h[a] > F->soff_h[b + 1] - F->soff_h[b];
  });
  std::vector<int2> tiles;
  for (int s : ord)
    for (int t0 = 0; t0 < F->soff_h[s + 1] - F->soff_h[s]; t0 += HB_T) tiles.push_back(make_int2(s, t0));
  F->ntile = (int)tiles.size();
  F->it_lap.assign(nsim, 0);
  F->ie_lap.assign(nsim, 0);
  F->it_st.assign(nsim, 0);
  F->ie_st.assign(nsim, 0);
  int rc;
  const size_t n2 = 2 * (size_t)n;
  if ((rc = dalloc(&F->soff, nsim + 1)) || (rc = dalloc(&F->body_of, n)) || (rc = dalloc(&F->tiles, tiles.size())) ||
      (rc = dalloc(&F->h, k)) || (rc = dalloc(&F->cap_b, k)) || (rc = dalloc(&F->e0_b, k)) ||
      (rc = dalloc(&F->alpha_b, k)) || (rc = dalloc(&F->x, n)) || (rc = dalloc(&F->y, n)) ||
      (rc = dalloc(&F->tha, n)) || (rc = dalloc(&F->dsda, n)) || (rc = dalloc(&F->rkappa, n)) ||
      (rc = dalloc(&F->area, k)) || (rc = dalloc(&F->xls, 2)) || (rc = dalloc(&F->dz, n)) || (rc = dalloc(&F->zk, k)) ||
      (rc = dalloc(&F->ckern, n)) || (rc = dalloc(&F->wdz, n)) || (rc = dalloc(&F->dens, n)) ||
      (rc = dalloc(&F->cf, n)) || (rc = dalloc(&F->uf, n2)) || (rc = dalloc(&F->wf, n2)) || (rc = dalloc(&F->cf2, n2)) ||
      (rc = dalloc(&F->x2, n2)) || (rc = dalloc(&F->y2, n2)) || (rc = dalloc(&F->rhs_lap, n)) ||
      (rc = dalloc(&F->ulap, n)) || (rc = dalloc(&F->dphi, n)) || (rc = dalloc(&F->rhs_st, n2)) ||
      (rc = dalloc(&F->pk, k)) || (rc = dalloc(&F->un, n)) || (rc = dalloc(&F->u, n)) || (rc = dalloc(&F->za, k)) ||
      (rc = dalloc(&F->udot, n)) || (rc = dalloc(&F->gw, n))) {
    free_sf(F);
    return rc;
  }
  if (!upload(F->soff, F->soff_h, F->st) || !upload(F->tiles, tiles, F->st) || !upload(F->h, h, F->st) ||
      !upload(F->cap_b, capv, F->st) || !upload(F->e0_b, e0v, F->st) || !upload(F->alpha_b, alv, F->st)) {
    free_sf(F);
    return (eb_set_error("eb_sweep_field_create: upload failed"), EB_ERR_HIP);
  }
  hipLaunchKernelGGL(k_body_of_f, dim3(k), dim3(256), 0, F->st, B->d_off, k, F->body_of);
  hipLaunchKernelGGL(k_pv_weights, dim3(k), dim3(256), 0, F->st, view(B), F->wf);
  if ((rc = eb_gmres_batch_create(&F->glap, nsim, F->soff_h.data(), 100, stream)) != EB_OK ||
      (rc = eb_gmres_batch_create(&F->gst, nsim, F->uoff_h.data(), 100, stream)) != EB_OK) {
    free_sf(F);
    return rc;
  }
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(F->st) != hipSuccess) {
    free_sf(F);
    return (eb_set_error("eb_sweep_field_create: kernel failed"), EB_ERR_HIP);
  }
  *out = F;
  return EB_OK;
}

extern "C" int eb_sweep_field_destroy(eb_sweep_field F) {
  free_sf(F);
  return EB_OK;
}

extern "C" int eb_sweep_field_set_mode(eb_sweep_field F, int32_t mode) {
  if (!F || (mode != EB_MODE_AUTO && mode != EB_MODE_FMM && mode != EB_MODE_DIRECT))
    return (eb_set_error("eb_sweep_field_set_mode: bad arguments"), EB_ERR_ARG);
  F->mode = mode;
  if (F->gexec) {  // the captured step holds the former mode's launches
    EB_HIP_CHECK(hipGraphExecDestroy(F->gexec));
    F->gexec = nullptr;
  }
  return F->sw ? eb_sweep_set_mode(F->sw, mode) : EB_OK;
}

// the segmented plan of one Hilbert product (created on first use, re-binned after)
static int sf_plan(eb_sweep_field F, eb_fmm_plan* P, int npts, const double* x, const double* y, const double2* w,
                   const std::vector<int>& off, int nmax, int nterms) {
  if (!*P)
    return eb_fmm_plan_create_segmented(P, npts, x, y, (const double*)w, F->nsim, off.data(),
                                        eb::seg_levmax(nmax, F->nsim), nterms, EB_VARIANT_HILBERT, (void*)F->st);
  return eb_fmm_plan_rebuild(*P, x, y, (const double*)w, (void*)F->st);
}

static int sf_lap_apply(void* ctx, const double* u, double* w, void* stream) {
  (void)stream;
  eb_sweep_field F = (eb_sweep_field)ctx;
  const int n = F->n;
  hipLaunchKernelGGL(k_real_dens, dim3(ew_blocks(n)), dim3(EW), 0, F->st, n, u, F->dens);
  if (sf_fmm(F)) {
    if (int rc = eb_fmm_apply(F->slap, (const double*)F->dens, (double*)F->cf, (void*)F->st)) return rc;
  } else {
    hipLaunchKernelGGL(k_hilbert_seg<true>, dim3(F->ntile), dim3(HB_T), 0, F->st, F->tiles, F->soff, 1, 0, 0, F->x,
                       F->y, F->dens, F->wdz, eb_gmres_batch_active(F->glap), F->cf);
  }
  hipLaunchKernelGGL(k_lap_point, dim3(ew_blocks(n)), dim3(EW), 0, F->st, n, F->body_of, F->h, u, F->cf, F->rkappa,
                     F->dsda, w);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

// F of every simulation.  budget = 0: the solves run the host-controlled batched DGMRES; budget
// > 0: the device-controlled one with at most `budget` Arnoldi steps and no host
// synchronisation anywhere (the graph-capturable form; geometry must exist already).
static int sweep_rhs(eb_sweep_field F, const double* rku, double* rkup, int budget) {
  eb_bodies B = F->B;
  const int n = F->n;
  void* s = (void*)F->st;
  int rc = eb_extractdat(B, rku, F->x, F->y, (double*)F->zk, (double*)F->dz, F->tha, F->dsda, F->rkappa,
                         (double*)F->ckern, F->area, nullptr, s);
  if (rc != EB_OK) return rc;
  hipLaunchKernelGGL(k_wdz_f, dim3(ew_blocks(n)), dim3(EW), 0, F->st, n, F->dz, F->h, F->body_of, F->wdz);
  EB_HIP_CHECK(hipGetLastError());
  if (!F->have_geom) {
    if ((rc = eb_sweep_create(&F->sw, B, F->nsim, F->sim_off.data(), F->x, F->y, (const double*)F->dz, F->dsda,
                              F->rkappa, (const double*)F->zk, s)) != EB_OK ||
        (rc = eb_sweep_set_mode(F->sw, F->mode)) != EB_OK)
      return rc;
    F->have_geom = true;
  } else if ((rc = eb_sweep_rebuild(F->sw, F->x, F->y, (const double*)F->dz, F->dsda, F->rkappa, (const double*)F->zk,
                                    s)) != EB_OK) {
    return rc;
  }
  // TOL 1e-11 for the Laplace side's DAPIF2 (laplace.f:212): 37 terms
  if (sf_fmm(F) && (rc = sf_plan(F, &F->slap, n, F->x, F->y, F->wdz, F->soff_h, F->maxn, 37)) != EB_OK) return rc;
  // GETBCs + SOLVE_LAPL of every simulation
  hipLaunchKernelGGL(k_getbcs, dim3(ew_blocks(n)), dim3(EW), 0, F->st, n, F->x, F->rhs_lap, F->ulap);
  EB_HIP_CHECK(hipGetLastError());
  std::vector<double> err(F->nsim);
  if (budget > 0)
    rc = eb_gmres_batch_solve_async(F->glap, sf_lap_apply, F, F->rhs_lap, F->ulap, 1.0e-10, budget);
  else
    rc = eb_gmres_batch_solve(F->glap, sf_lap_apply, F, F->rhs_lap, F->ulap, 1.0e-10, F->it_lap.data(), err.data(),
                              F->ie_lap.data());
  if (rc != EB_OK) return rc;
  // PVINTEV + DIRICH_NEU_MAP
  hipLaunchKernelGGL(k_pv_interp, dim3(B->nbody), dim3(SP_THREADS), lds_bytes(B, 2), F->st, view(B), F->x, F->y,
                     F->ulap, F->dz, F->x2, F->y2, F->uf);
  if (sf_fmm(F)) {
    // all 2N points of every simulation, the midpoints weighted h and the original points 0; the
    // original points' values are the ones read (TOL 1e-10, pv_integral.f:175: 33 terms)
    if ((rc = sf_plan(F, &F->spv, 2 * n, F->x2, F->y2, F->wf, F->soff2_h, 2 * F->maxn, 33)) != EB_OK ||
        (rc = eb_fmm_apply(F->spv, (const double*)F->uf, (double*)F->cf2, s)) != EB_OK)
      return rc;
  } else {
    hipLaunchKernelGGL(k_hilbert_seg<false>, dim3(F->ntile), dim3(HB_T), 0, F->st, F->tiles, F->soff, 2, 0, 1, F->x2,
                       F->y2, F->uf, F->wf, (const int*)nullptr, F->cf2);
  }
  hipLaunchKernelGGL(k_dn_post, dim3(B->nbody), dim3(SP_THREADS), lds_bytes(B, 1), F->st, view(B), F->x, F->y, F->ulap,
                     F->cf2, F->dsda, nullptr, nullptr, F->dphi);
  // SETVELF
  hipLaunchKernelGGL(k_setvelf, dim3(B->nbody), dim3(SP_THREADS), lds_bytes(B, 2), F->st, view(B), F->x, F->y, F->dz,
                     F->dsda, F->dphi, F->cap_b, F->e0_b, F->alpha_b, (double2*)F->rhs_st);
  EB_HIP_CHECK(hipGetLastError());
  // SOLVEI of every simulation
  if (budget > 0)
    rc = eb_sweep_solve_async(F->sw, F->gst, F->rhs_st, (double*)F->u, budget, s);
  else
    rc = eb_sweep_solve(F->sw, F->gst, F->rhs_st, (double*)F->u, nullptr, nullptr, F->it_st.data(), err.data(),
                        F->ie_st.data(), s);
  if (rc != EB_OK) return rc;
  // INTERFACENEW: FDIFFF of u, CADAPMOD of every simulation, corrections
  if ((rc = eb_fdifff(B, (const double*)F->u, (double*)F->udot, s)) != EB_OK) return rc;
  if ((rc = eb_sweep_layer(F->sw, (const double*)F->u, (double*)F->gw, EB_VARIANT_MOD, s)) != EB_OK) return rc;
  hipLaunchKernelGGL(k_interface_point, dim3(ew_blocks(n)), dim3(EW), 0, F->st, n, F->body_of, F->h, F->x, F->y, F->dz,
                     F->dsda, F->rkappa, F->ckern, F->u, F->udot, F->gw, F->cap_b, F->un);
  EB_HIP_CHECK(hipGetLastError());
  return eb_ftail(B, F->tha, F->un, (const double*)F->dz, F->dsda, rkup, s);
}

extern "C" int eb_sweep_field_rhs(eb_sweep_field F, const double* rku, double* rkup) {
  if (!F || !rku || !rkup) return (eb_set_error("eb_sweep_field_rhs: bad arguments"), EB_ERR_ARG);
  return sweep_rhs(F, rku, rkup, 0);
}

// One Heun step (electric.f:773-783) on the field's buffers: phi0 = dt F(ybar), y = ybar + phi0,
// ybar += phi0/2 + dt F(y)/2, with the RK2 kernels' roundings.
static int sweep_rk2_step(eb_sweep_field F, double* ybar, double dt, int budget) {
  const int neq = 2 * F->n;
  void* s = (void*)F->st;
  int rc;
  if ((rc = sweep_rhs(F, ybar, F->dy, budget)) != EB_OK) return rc;
  if ((rc = eb_rk2_stage1(neq, dt, ybar, F->dy, F->phi0, F->yst, s)) != EB_OK) return rc;
  if ((rc = sweep_rhs(F, F->yst, F->dy, budget)) != EB_OK) return rc;
  return eb_rk2_stage2(neq, dt, F->phi0, F->dy, ybar, s);
}

// One classic RK4 step of every simulation (phi0: the slopes' weighted sum, yst: the stage states)
static int sweep_rk4_step(eb_sweep_field F, double* ybar, double dt, int budget) {
  const int neq = 2 * F->n;
  void* s = (void*)F->st;
  int rc;
  for (int q = 0; q < 4; ++q) {
    if ((rc = sweep_rhs(F, q == 0 ? ybar : F->yst, F->dy, budget)) != EB_OK) return rc;
    if ((rc = eb_rk4_stage(neq, q, dt, ybar, F->dy, F->phi0, F->yst, s)) != EB_OK) return rc;
  }
  return EB_OK;
}

static int sweep_step(eb_sweep_field F, double* ybar, double dt, int budget, int order) {
  return order == 4 ? sweep_rk4_step(F, ybar, dt, budget) : sweep_rk2_step(F, ybar, dt, budget);
}

static int sweep_step_graph(eb_sweep_field F, double* ybar, double dt, int32_t budget, int32_t* redone, int order,
                            const char* name) {
  if (!F || !ybar || budget < 1) return (eb_set_error("%s: bad arguments", name), EB_ERR_ARG);
  if (F->st == nullptr)
    return (eb_set_error("%s: the field must use a non-default stream (graph capture)", name), EB_ERR_ARG);
  const size_t neq = 2 * (size_t)F->n;
  int rc;
  if (!F->ysave) {
    if ((rc = dalloc(&F->ysave, neq)) || (rc = dalloc(&F->yst, neq)) || (rc = dalloc(&F->phi0, neq)) ||
        (rc = dalloc(&F->dy, neq)))
      return rc;
  }
  if (!F->have_geom && (rc = sweep_rhs(F, ybar, F->dy, 0)) != EB_OK) return rc;  // creates the sweep operator
  if (!F->gexec || F->g_ybar != ybar || F->g_dt != dt || F->g_budget != budget || F->g_order != order) {
    if (F->gexec) {
      EB_HIP_CHECK(hipGraphExecDestroy(F->gexec));
      F->gexec = nullptr;
    }
    EB_HIP_CHECK(hipStreamSynchronize(F->st));
    hipGraph_t graph = nullptr;
    EB_HIP_CHECK(hipStreamBeginCapture(F->st, hipStreamCaptureModeThreadLocal));
    int crc = hipMemcpyAsync(F->ysave, ybar, sizeof(double) * neq, hipMemcpyDeviceToDevice, F->st) == hipSuccess
                  ? EB_OK
                  : EB_ERR_HIP;
    if (crc == EB_OK) crc = eb_gmres_batch_clear_fail(F->glap);
    if (crc == EB_OK) crc = eb_gmres_batch_clear_fail(F->gst);
    if (crc == EB_OK) crc = sweep_step(F, ybar, dt, budget, order);
    const hipError_t e = hipStreamEndCapture(F->st, &graph);
    if (crc != EB_OK || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      if (crc == EB_OK) eb_set_error("%s: capture failed: %s", name, hipGetErrorString(e));
      return crc != EB_OK ? crc : EB_ERR_HIP;
    }
    const hipError_t ie = hipGraphInstantiate(&F->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
      F->gexec = nullptr;
      return (eb_set_error("%s: instantiate failed: %s", name, hipGetErrorString(ie)), EB_ERR_HIP);
    }
    F->g_ybar = ybar;
    F->g_dt = dt;
    F->g_budget = budget;
    F->g_order = order;
  }
  EB_HIP_CHECK(hipGraphLaunch(F->gexec, F->st));
  int32_t fl = 0, fs = 0;
  if ((rc = eb_gmres_batch_results(F->glap, F->it_lap.data(), nullptr, F->ie_lap.data(), &fl)) != EB_OK ||
      (rc = eb_gmres_batch_results(F->gst, F->it_st.data(), nullptr, F->ie_st.data(), &fs)) != EB_OK)
    return rc;
  if (redone) *redone = 0;
  if (fl + fs > 0) {
    // some system needed more than `budget` Arnoldi steps: the step again from the saved state
    // with the host-controlled solves
    EB_HIP_CHECK(hipMemcpyAsync(ybar, F->ysave, sizeof(double) * neq, hipMemcpyDeviceToDevice, F->st));
    if ((rc = sweep_step(F, ybar, dt, 0, order)) != EB_OK) return rc;
    EB_HIP_CHECK(hipStreamSynchronize(F->st));
    if (redone) *redone = 1;
  }
  return EB_OK;
}

extern "C" int eb_sweep_field_rk2_graph(eb_sweep_field F, double* ybar, double dt, int32_t budget, int32_t* redone) {
  return sweep_step_graph(F, ybar, dt, budget, redone, 2, "eb_sweep_field_rk2_graph");
}

extern "C" int eb_sweep_field_rk4_graph(eb_sweep_field F, double* ybar, double dt, int32_t budget, int32_t* redone) {
  return sweep_step_graph(F, ybar, dt, budget, redone, 4, "eb_sweep_field_rk4_graph");
}

extern "C" int eb_sweep_field_iterations(eb_sweep_field F, int32_t* lap_iter, int32_t* lap_ierr, int32_t* stokes_iter,
                                         int32_t* stokes_ierr) {
  if (!F) return (eb_set_error("eb_sweep_field_iterations: bad arguments"), EB_ERR_ARG);
  for (int s = 0; s < F->nsim; ++s) {
    if (lap_iter) lap_iter[s] = F->it_lap[s];
    if (lap_ierr) lap_ierr[s] = F->ie_lap[s];
    if (stokes_iter) stokes_iter[s] = F->it_st[s];
    if (stokes_ierr) stokes_ierr[s] = F->ie_st[s];
  }
  return EB_OK;
}
