// eb_spectral.hip -- per-body spectral geometry, RK2 stage updates and the Stokes matvec
// epilogue, batched over all bodies (drops) of one or many simulations.
//
// One 512-thread workgroup owns one body; the body's sequence lives in LDS for the whole
// transform chain (FFT, multiply, inverse FFT), so each kernel reads its inputs from HBM once
// and writes its outputs once.  The FFT (eb_fft.h) runs in place in LDS: Stockham stages of
// radix 2, 3, 4, 5 and 8 in registers, direct sums for other small primes, and Bluestein's
// method for a large prime factor (body lengths rounded to even give e.g. 1942 = 2 x 971).
// Workgroups take the bodies in order of decreasing cost, so the longest transforms start
// first.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

#include "../../include/eb_spectral.h"
#include "eb_bodies.h"
#include "eb_fft_plan.h"

using namespace eb;

// ------------------------------------------------------------------------------------------
// kernels (one block per body; dynamic LDS holds nseq sequences of nmax entries followed by
// the Bluestein workspace)
// ------------------------------------------------------------------------------------------

__global__ __launch_bounds__(SP_THREADS) void k_dcfft(BodyView B, double2* __restrict__ c, int sign) {
  extern __shared__ double2 lds[];
  const BodyCtx bc = body_ctx(B);
  double2* a = lds;
  for (int i = threadIdx.x; i < bc.N; i += blockDim.x) a[i] = c[bc.o0 + i];
  __syncthreads();
  fft(a, lds + B.nmax, bc, B, sign);
  for (int i = threadIdx.x; i < bc.N; i += blockDim.x) c[bc.o0 + i] = a[i];
}

__global__ __launch_bounds__(SP_THREADS) void k_fdifff(BodyView B, const double2* __restrict__ f,
                                                        double2* __restrict__ df) {
  extern __shared__ double2 lds[];
  const BodyCtx bc = body_ctx(B);
  const int N = bc.N;
  double2* a = lds;
  double2* ws = lds + B.nmax;
  for (int i = threadIdx.x; i < N; i += blockDim.x) a[i] = f[bc.o0 + i];
  __syncthreads();
  fft(a, ws, bc, B, -1);
  const double invn = 1.0 / N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const double w = fdiff_k(i, N) * invn;  // (i k / N) * F
    const double2 v = a[i];
    a[i] = make_double2(-w * v.y, w * v.x);
  }
  __syncthreads();
  fft(a, ws, bc, B, +1);
  for (int i = threadIdx.x; i < N; i += blockDim.x) df[bc.o0 + i] = a[i];
}

__global__ __launch_bounds__(SP_THREADS) void k_filter(BodyView B, double2* __restrict__ c) {
  extern __shared__ double2 lds[];
  const BodyCtx bc = body_ctx(B);
  for (int i = threadIdx.x; i < bc.N; i += blockDim.x) lds[i] = c[bc.o0 + i];
  __syncthreads();
  filter_block(lds, bc.N);
  for (int i = threadIdx.x; i < bc.N; i += blockDim.x) c[bc.o0 + i] = lds[i];
}

// EQUIT + FOURINT (electric.f:1270-1331) of one body: a and c are LDS sequences, ws the
// transform workspace.  T(i) is left in c[i].x; returns synchronised.
__device__ __forceinline__ void equit_block(const BodyCtx& bc, const BodyView& B, const double* __restrict__ tha,
                            const double* __restrict__ un, double2* a, double2* c, double2* ws) {
  const int N = bc.N, o0 = bc.o0, n = N / 2;
  for (int i = threadIdx.x; i < N; i += blockDim.x) a[i] = make_double2(-tha[o0 + i] * un[o0 + i], 0.0);
  __syncthreads();
  fft(a, ws, bc, B, -1);
  const double invn = 1.0 / N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    double2 v = a[i];
    v = make_double2(v.x * invn, v.y * invn);
    if (i == n) v = make_double2(0, 0);  // zf(n+1) = 0
    a[i] = v;
    // FOURINT: zfi(i) = zf(i)/(i*eye) for 1 <= i <= n-1, the mirrored ones with -, zero elsewhere
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
  const double2 zf0 = a[0];
  const double dadt = zf0.x;
  const double dalph = 2.0 * 3.14159265358979323846 / N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const double alpha = dalph * (double)i;
    c[i].x = -(zf0.x * alpha + c[i].x) + alpha * dadt;
  }
  __syncthreads();
}

__global__ __launch_bounds__(SP_THREADS) void k_equit(BodyView B, const double* __restrict__ tha,
                                                       const double* __restrict__ un, double* __restrict__ t) {
  extern __shared__ double2 lds[];
  const BodyCtx bc = body_ctx(B);
  double2* a = lds;
  double2* c = lds + B.nmax;
  equit_block(bc, B, tha, un, a, c, lds + 2 * B.nmax);
  for (int i = threadIdx.x; i < bc.N; i += blockDim.x) t[bc.o0 + i] = c[i].x;
}

__global__ __launch_bounds__(SP_THREADS) void k_ftail(BodyView B, const double* __restrict__ tha,
                                                       const double* __restrict__ un, const double2* __restrict__ dz,
                                                       const double* __restrict__ dsda, double2* __restrict__ rkup) {
  extern __shared__ double2 lds[];
  const BodyCtx bc = body_ctx(B);
  const int N = bc.N, o0 = bc.o0;
  double2* a = lds;
  double2* c = lds + B.nmax;
  double2* ws = lds + 2 * B.nmax;
  equit_block(bc, B, tha, un, a, c, ws);
  // zf = (-Un i dz - T dz) / dsda  (electric.f:1084-1086), T from c[i].x
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const double2 d = dz[o0 + i];
    const double u = un[o0 + i], tt = c[i].x, s = dsda[o0 + i];
    a[i] = make_double2((u * d.y - tt * d.x) / s, (-u * d.x - tt * d.y) / s);
  }
  __syncthreads();
  fft(a, ws, bc, B, -1);
  const double invn = 1.0 / N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) a[i] = make_double2(a[i].x * invn, a[i].y * invn);
  __syncthreads();
  filter_block(a, N);
  fft(a, ws, bc, B, +1);
  for (int i = threadIdx.x; i < N; i += blockDim.x) rkup[o0 + i] = a[i];
}

// EXTRACTDAT (electric.f:636-716), one block per body.  One forward FFT of z gives both
// derivatives (the reference applies FDIFFF twice; the mean and Nyquist rules are the same).
__global__ __launch_bounds__(SP_THREADS) void k_extractdat(BodyView B, const double2* __restrict__ z,
                                                            double* __restrict__ x, double* __restrict__ y,
                                                            double2* __restrict__ zk, double2* __restrict__ dz,
                                                            double* __restrict__ tha, double* __restrict__ dsda,
                                                            double* __restrict__ rkappa, double2* __restrict__ ckern,
                                                            double* __restrict__ area, double* __restrict__ red4) {
  extern __shared__ double2 lds[];
  __shared__ double red[SP_THREADS];
  const BodyCtx bc = body_ctx(B);
  const int k = bc.k, N = bc.N, o0 = bc.o0;
  double2* a = lds;
  double2* c = lds + B.nmax;
  double2* ws = lds + 2 * B.nmax;
  for (int i = threadIdx.x; i < N; i += blockDim.x) a[i] = z[o0 + i];
  __syncthreads();
  fft(a, ws, bc, B, -1);
  const double invn = 1.0 / N;
  if (threadIdx.x == 0 && zk) zk[k] = make_double2(a[0].x * invn, a[0].y * invn);
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const double kk = fdiff_k(i, N);
    const double2 v = make_double2(a[i].x * invn, a[i].y * invn);
    c[i] = make_double2(-kk * kk * v.x, -kk * kk * v.y);  // (ik)^2 F / N
    a[i] = make_double2(-kk * v.y, kk * v.x);             // (ik) F / N
  }
  __syncthreads();
  fft(a, ws, bc, B, +1);  // z_alpha
  fft(c, ws, bc, B, +1);  // z_alpha_alpha
  double sarea = 0.0, xmn = 1e300, xmx = -1e300, ymn = 1e300, ymx = -1e300;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const int g = o0 + i;
    const double2 zz = z[g];
    const double xdot = a[i].x, ydot = a[i].y, xddot = c[i].x, yddot = c[i].y;
    const double dsdt = sqrt(xdot * xdot + ydot * ydot);
    const double kap = -(xdot * yddot - ydot * xddot) / (dsdt * dsdt * dsdt);
    if (x) x[g] = zz.x;
    if (y) y[g] = zz.y;
    if (rkappa) rkappa[g] = kap;
    if (ckern) {
      // -(xdd + i ydd) / (xd + i yd)
      const double den = xdot * xdot + ydot * ydot;
      ckern[g] = make_double2(-(xddot * xdot + yddot * ydot) / den, -(yddot * xdot - xddot * ydot) / den);
    }
    if (tha) tha[g] = -kap * dsdt;
    if (dz) dz[g] = make_double2(-xdot, -ydot);
    if (dsda) dsda[g] = hypot(xdot, ydot);
    sarea += zz.y * (-xdot);
    xmn = fmin(xmn, zz.x);
    xmx = fmax(xmx, zz.x);
    ymn = fmin(ymn, zz.y);
    ymx = fmax(ymx, zz.y);
  }
  const double s = block_sum(sarea, red);
  if (threadIdx.x == 0 && area) area[k] = 2.0 * 3.14159265358979323846 * (s * invn);
  // min / max through the same LDS buffer
  red[threadIdx.x] = xmn;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) red4[4 * k + 0] = red[0];
  __syncthreads();
  red[threadIdx.x] = xmx;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) red4[4 * k + 1] = red[0];
  __syncthreads();
  red[threadIdx.x] = ymn;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) red4[4 * k + 2] = red[0];
  __syncthreads();
  red[threadIdx.x] = ymx;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) red4[4 * k + 3] = red[0];
}

// xleft = min x, slength = max(xmax - xmin, ymax) over all bodies (electric.f:710-711)
__global__ void k_extract_final(const double* __restrict__ red4, int nbody, double* __restrict__ out2) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double xmn = 1e10, xmx = -1e10, ymx = -1e10;
  for (int k = 0; k < nbody; ++k) {
    xmn = fmin(xmn, red4[4 * k]);
    xmx = fmax(xmx, red4[4 * k + 1]);
    ymx = fmax(ymx, red4[4 * k + 3]);
  }
  out2[0] = xmn;
  out2[1] = fmax(xmx - xmn, ymx);
}

__global__ void k_rk2_1(int neq, double dt, const double* __restrict__ ybar, const double* __restrict__ dy,
                        double* __restrict__ phi0, double* __restrict__ y) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= neq) return;
  // separate roundings, as the reference's two statements (no fused multiply-add)
  const double p = dt * dy[i];
  phi0[i] = p;
  y[i] = ybar[i] + p;
}

__global__ void k_rk2_2(int neq, double dt, const double* __restrict__ phi0, const double* __restrict__ dy,
                        double* __restrict__ ybar) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= neq) return;
  const double p1 = dt * dy[i];
  ybar[i] = ybar[i] + 0.5 * phi0[i] + 0.5 * p1;
}

// Classic fourth-order Runge-Kutta over neq doubles in four stage updates, each after one
// evaluation dy = F(stage state): acc sums the weighted slopes, ys is the next stage's state.
//   stage 0: acc = dy,       ys = y + (dt/2) dy
//   stage 1: acc += 2 dy,    ys = y + (dt/2) dy
//   stage 2: acc += 2 dy,    ys = y + dt dy
//   stage 3: y += (dt/6) (acc + dy)
__global__ void k_rk4(int neq, int stage, double dt, double* __restrict__ y, const double* __restrict__ dy,
                      double* __restrict__ acc, double* __restrict__ ys) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= neq) return;
  const double d = dy[i];
  switch (stage) {
    case 0:
      acc[i] = d;
      ys[i] = y[i] + (0.5 * dt) * d;
      break;
    case 1:
    case 2:
      acc[i] = acc[i] + 2.0 * d;
      ys[i] = y[i] + (stage == 1 ? 0.5 * dt : dt) * d;
      break;
    default:
      y[i] = y[i] + (dt / 6.0) * (acc[i] + d);
      break;
  }
}

// The deformation D = (L - B) / (L + B) of every body: L and B are the semi-axes of the ellipse
// with the body's area and second moments of area, from trapezoid sums of Green's-theorem
// integrands with the spectral tangent (FDIFFF's wavenumbers), about the mean of the body's
// points (the central moments do not depend on the origin; taking it near the body keeps the
// subtraction A c^2 from cancelling).  One workgroup per body.
__global__ __launch_bounds__(SP_THREADS) void k_deformation(BodyView B, const double2* __restrict__ z,
                                                            double* __restrict__ D) {
  extern __shared__ double2 lds[];
  __shared__ double red[SP_THREADS];
  const BodyCtx bc = body_ctx(B);
  const int N = bc.N;
  double2* a = lds;                // the tangent
  double2* zz = lds + B.nmax;      // the points, about their mean
  double2* ws = lds + 2 * B.nmax;
  double sx = 0.0, sy = 0.0;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const double2 v = z[bc.o0 + i];
    a[i] = v;
    sx += v.x;
    sy += v.y;
  }
  const double x0 = block_sum(sx, red) / N, y0 = block_sum(sy, red) / N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) zz[i] = make_double2(a[i].x - x0, a[i].y - y0);
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
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const double x = zz[i].x, y = zz[i].y, xa = a[i].x, ya = a[i].y;
    s[0] += x * ya - y * xa;  // 2 A
    s[1] += x * x * ya;       // 2 mx
    s[2] += y * y * xa;       // -2 my
    s[3] += x * x * x * ya;   // 3 ixx
    s[4] += y * y * y * xa;   // -3 iyy
    s[5] += x * x * y * ya;   // 2 ixy
  }
  double t[6];
  for (int q = 0; q < 6; ++q) t[q] = block_sum(s[q], red);
  if (threadIdx.x == 0) {
    const double h = 2.0 * 3.14159265358979323846 / N;
    const double A = 0.5 * h * t[0], mx = 0.5 * h * t[1], my = -0.5 * h * t[2];
    const double ixx = h * t[3] / 3.0, iyy = -h * t[4] / 3.0, ixy = 0.5 * h * t[5];
    const double cx = mx / A, cy = my / A;
    const double p = ixx - A * cx * cx, c = iyy - A * cy * cy, b = ixy - A * cx * cy;
    const double rt = sqrt(0.25 * (p - c) * (p - c) + b * b);
    const double sp = sqrt(0.5 * (p + c) + rt), sm = sqrt(0.5 * (p + c) - rt);
    D[bc.k] = (sp - sm) / (sp + sm);
  }
}

// ------------------------------------------------------------------------------------------
// FASMVP epilogue (stokesbubb.f:527-549, 585-611)
// ------------------------------------------------------------------------------------------

// per body: za = h sum dens, pk = (2h/pi) sum Re(dens dz / (z - zk)^2)
__global__ __launch_bounds__(SP_THREADS) void k_fasmvp_sums(BodyView B, const double* __restrict__ x,
                                                             const double* __restrict__ y, const double2* __restrict__ dz,
                                                             const double2* __restrict__ zk, const double* __restrict__ h,
                                                             const double2* __restrict__ dens, double2* __restrict__ za,
                                                             double* __restrict__ pk) {
  __shared__ double red[SP_THREADS];
  const int k = blockIdx.x, N = B.nd[k], o0 = B.off[k];
  const double2 c = zk[k];
  double sr = 0, si = 0, sp = 0;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const int g = o0 + i;
    const double2 d = dens[g];
    sr += d.x;
    si += d.y;
    const double2 zd = make_double2(x[g] - c.x, y[g] - c.y);
    const double2 zd2 = cmul(zd, zd);
    const double2 num = cmul(d, dz[g]);
    const double den = zd2.x * zd2.x + zd2.y * zd2.y;
    sp += (num.x * zd2.x + num.y * zd2.y) / den;  // Re(num / zd2)
  }
  const double tr = block_sum(sr, red), ti = block_sum(si, red), tp = block_sum(sp, red);
  if (threadIdx.x == 0) {
    const double hk = h[k];
    za[k] = make_double2(hk * tr, hk * ti);
    pk[k] = 2.0 * hk * tp / 3.14159265358979323846;
  }
}

// w = dens + za + GW + self - (i/2)(pinf - pk_k) z, with pinf = -pk_1 and pk_k shifted by pinf
__global__ void k_fasmvp_point(BodyView B, int nbody, int n, const double* __restrict__ x, const double* __restrict__ y,
                               const double2* __restrict__ dz, const double* __restrict__ dsdth,
                               const double* __restrict__ rkappa, const double* __restrict__ h,
                               const double2* __restrict__ dens, const double2* __restrict__ za,
                               const double* __restrict__ pk, double2* __restrict__ gw, double2* __restrict__ w,
                               const int* __restrict__ body_of) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const int k = body_of[g];
  const double pi = 3.14159265358979323846;
  const double pinf = -pk[0];
  const double pkk = pk[k] + pinf;
  const double2 zu = dens[g];
  const double2 d = dz[g];
  const double s = dsdth[g];
  const double2 zn = make_double2(d.y / s, -d.x / s);  // -i dz / dsdth
  const double qwi = h[k];
  const double c0 = qwi * s * 0.5 * rkappa[g] / pi;
  const double2 zn2 = cmul(zn, zn);
  // zself = zu c0 - conj(zu) c0 zn^2
  const double2 t = cmul(make_double2(zu.x, -zu.y), zn2);
  const double2 zself = make_double2(c0 * (zu.x - t.x), c0 * (zu.y - t.y));
  // -0.5 i (pinf - pkk) z
  const double a = -0.5 * (pinf - pkk);
  const double2 pr = make_double2(-a * y[g], a * x[g]);
  const double2 st = make_double2(gw[g].x + zself.x + pr.x, gw[g].y + zself.y + pr.y);
  gw[g] = st;
  const double2 zak = za[k];
  w[g] = make_double2(zu.x + zak.x + st.x, zu.y + zak.y + st.y);
}

__global__ void k_body_of(const int* __restrict__ off, int nbody, int* __restrict__ body_of) {
  const int k = blockIdx.x;
  if (k >= nbody) return;
  for (int i = off[k] + threadIdx.x; i < off[k + 1]; i += blockDim.x) body_of[i] = k;
}

__global__ void k_wdz(int n, const double2* __restrict__ dz, const double* __restrict__ h,
                      const int* __restrict__ body_of, double2* __restrict__ wdz) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const double hk = h[body_of[g]];
  wdz[g] = make_double2(dz[g].x * hk, dz[g].y * hk);
}

__global__ void k_shift_pk(int nbody, double* __restrict__ pk_out, const double* __restrict__ pk) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nbody) return;
  pk_out[k] = pk[k] - pk[0];
}

// ------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------

static void free_bodies(eb_bodies B) {
  if (!B) return;
  void* ptrs[] = {B->d_nd, B->d_off, B->d_plan_of, B->d_order, B->d_plans, B->d_stages, B->d_tab, B->d_red};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  delete B;
}


extern "C" int eb_bodies_create(eb_bodies* out, int32_t nbody, const int32_t* nd, void* stream) {
  if (!out || nbody < 1 || !nd) {
    eb_set_error("eb_bodies_create: bad arguments");
    return EB_ERR_ARG;
  }
  *out = nullptr;
  eb_bodies B = new eb_bodies_s();
  B->nbody = nbody;
  B->nd.assign(nd, nd + nbody);
  B->off.assign(nbody + 1, 0);
  for (int k = 0; k < nbody; ++k) {
    if (nd[k] < 2 || nd[k] > EB_SPECTRAL_NMAX) {
      eb_set_error("eb_bodies_create: body %d has %d points (2..%d supported)", k, nd[k], EB_SPECTRAL_NMAX);
      free_bodies(B);
      return EB_ERR_ARG;
    }
    B->off[k + 1] = B->off[k] + nd[k];
    B->nmax = std::max(B->nmax, (int)nd[k]);
  }
  B->n = B->off[nbody];
  // LDS of a workgroup: two sequences of nmax entries, the Bluestein workspace and the
  // static reduction buffer of SP_THREADS doubles
  int dev = 0, maxlds = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) {
    eb_set_error("eb_bodies_create: no device");
    free_bodies(B);
    return EB_ERR_HIP;
  }
  const long long avail = ((long long)maxlds - (long long)sizeof(double) * SP_THREADS) / (long long)sizeof(double2);
  if (avail < 2LL * B->nmax) {
    eb_set_error("eb_bodies_create: a body of %d points needs %zu B of LDS, the device allows %d", B->nmax,
                 sizeof(double2) * 2 * (size_t)B->nmax + sizeof(double) * SP_THREADS, maxlds);
    free_bodies(B);
    return EB_ERR_ARG;
  }
  const int ext_budget = (int)std::min<long long>(avail - 2LL * B->nmax, ebfft::EMAX);
  ebfft::Tables T;
  std::vector<int> plan_of(nbody), order(nbody);
  for (int k = 0; k < nbody; ++k) plan_of[k] = ebfft::plan_for(T, nd[k], ext_budget);
  for (const ebfft::Plan& p : T.plans)
    if (p.P) B->ext = std::max(B->ext, p.Ns * p.M);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int i, int j) { return T.plans[plan_of[i]].cost > T.plans[plan_of[j]].cost; });
  hipStream_t st = (hipStream_t)stream;
  int rc;
  if ((rc = dalloc(&B->d_nd, nbody)) || (rc = dalloc(&B->d_off, nbody + 1)) || (rc = dalloc(&B->d_plan_of, nbody)) ||
      (rc = dalloc(&B->d_order, nbody)) || (rc = dalloc(&B->d_plans, T.plans.size())) ||
      (rc = dalloc(&B->d_stages, T.stages.size())) || (rc = dalloc(&B->d_tab, T.tab.size())) ||
      (rc = dalloc(&B->d_red, 4 * (size_t)nbody))) {
    free_bodies(B);
    return rc;
  }
  const bool ok = upload(B->d_nd, B->nd, st) && upload(B->d_off, B->off, st) && upload(B->d_plan_of, plan_of, st) &&
                  upload(B->d_order, order, st) && upload(B->d_plans, T.plans, st) &&
                  upload(B->d_stages, T.stages, st) && upload(B->d_tab, T.tab, st) &&
                  hipStreamSynchronize(st) == hipSuccess;
  if (!ok) {
    eb_set_error("eb_bodies_create: upload failed");
    free_bodies(B);
    return EB_ERR_HIP;
  }
  *out = B;
  return EB_OK;
}

extern "C" int eb_bodies_destroy(eb_bodies B) {
  free_bodies(B);
  return EB_OK;
}

extern "C" int eb_bodies_npoints(eb_bodies B) { return B ? B->n : 0; }


extern "C" int eb_dcfft(eb_bodies B, double* c, int32_t sign, void* stream) {
  if (!B || !c || (sign != 1 && sign != -1)) return (eb_set_error("eb_dcfft: bad arguments"), EB_ERR_ARG);
  hipLaunchKernelGGL(k_dcfft, dim3(B->nbody), dim3(SP_THREADS), lds_bytes(B, 1), (hipStream_t)stream, view(B),
                     (double2*)c, sign);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" int eb_fdifff(eb_bodies B, const double* f, double* df, void* stream) {
  if (!B || !f || !df) return (eb_set_error("eb_fdifff: bad arguments"), EB_ERR_ARG);
  hipLaunchKernelGGL(k_fdifff, dim3(B->nbody), dim3(SP_THREADS), lds_bytes(B, 1), (hipStream_t)stream, view(B),
                     (const double2*)f, (double2*)df);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" int eb_filter(eb_bodies B, double* cn, void* stream) {
  if (!B || !cn) return (eb_set_error("eb_filter: bad arguments"), EB_ERR_ARG);
  hipLaunchKernelGGL(k_filter, dim3(B->nbody), dim3(SP_THREADS), sizeof(double2) * (size_t)B->nmax, (hipStream_t)stream, view(B),
                     (double2*)cn);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" int eb_equit(eb_bodies B, const double* tha, const double* un, double* t, void* stream) {
  if (!B || !tha || !un || !t) return (eb_set_error("eb_equit: bad arguments"), EB_ERR_ARG);
  hipLaunchKernelGGL(k_equit, dim3(B->nbody), dim3(SP_THREADS), lds_bytes(B, 2), (hipStream_t)stream, view(B), tha,
                     un, t);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" int eb_ftail(eb_bodies B, const double* tha, const double* un, const double* dz, const double* dsda,
                        double* rkup, void* stream) {
  if (!B || !tha || !un || !dz || !dsda || !rkup) return (eb_set_error("eb_ftail: bad arguments"), EB_ERR_ARG);
  const size_t lds = lds_bytes(B, 2);
  hipLaunchKernelGGL(k_ftail, dim3(B->nbody), dim3(SP_THREADS), lds, (hipStream_t)stream, view(B), tha, un,
                     (const double2*)dz, dsda, (double2*)rkup);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" int eb_extractdat(eb_bodies B, const double* rku, double* x, double* y, double* zk, double* dz,
                             double* tha, double* dsda, double* rkappa, double* ckern, double* area,
                             double* xleft_slength, void* stream) {
  if (!B || !rku) return (eb_set_error("eb_extractdat: bad arguments"), EB_ERR_ARG);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_extractdat, dim3(B->nbody), dim3(SP_THREADS), lds_bytes(B, 2), st, view(B),
                     (const double2*)rku, x, y, (double2*)zk, (double2*)dz, tha, dsda, rkappa, (double2*)ckern,
                     area, B->d_red);
  if (xleft_slength) hipLaunchKernelGGL(k_extract_final, dim3(1), dim3(64), 0, st, B->d_red, B->nbody, xleft_slength);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" int eb_rk2_stage1(int32_t neq, double dt, const double* ybar, const double* dy, double* phi0, double* y,
                             void* stream) {
  if (neq < 1) return EB_OK;
  hipLaunchKernelGGL(k_rk2_1, dim3((neq + 255) / 256), dim3(256), 0, (hipStream_t)stream, neq, dt, ybar, dy, phi0, y);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" int eb_rk2_stage2(int32_t neq, double dt, const double* phi0, const double* dy, double* ybar,
                             void* stream) {
  if (neq < 1) return EB_OK;
  hipLaunchKernelGGL(k_rk2_2, dim3((neq + 255) / 256), dim3(256), 0, (hipStream_t)stream, neq, dt, phi0, dy, ybar);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" int eb_rk4_stage(int32_t neq, int32_t stage, double dt, double* y, const double* dy, double* acc, double* ys,
                            void* stream) {
  if (stage < 0 || stage > 3 || !y || !dy || !acc || (stage < 3 && !ys))
    return (eb_set_error("eb_rk4_stage: bad arguments"), EB_ERR_ARG);
  if (neq < 1) return EB_OK;
  hipLaunchKernelGGL(k_rk4, dim3((neq + 255) / 256), dim3(256), 0, (hipStream_t)stream, neq, stage, dt, y, dy, acc, ys);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" int eb_deformation(eb_bodies B, const double* rku, double* D, void* stream) {
  if (!B || !rku || !D) return (eb_set_error("eb_deformation: bad arguments"), EB_ERR_ARG);
  hipLaunchKernelGGL(k_deformation, dim3(B->nbody), dim3(SP_THREADS), lds_bytes(B, 2), (hipStream_t)stream, view(B),
                     (const double2*)rku, D);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

// ------------------------------------------------------------------------------------------
// matvec
// ------------------------------------------------------------------------------------------
struct eb_matvec_s {
  eb_bodies B = nullptr;
  eb_fmm_plan fmm = nullptr;
  int n = 0;
  double *x = nullptr, *y = nullptr, *dsdth = nullptr, *rkappa = nullptr, *h = nullptr, *pk = nullptr;
  double2 *dz = nullptr, *zk = nullptr, *wdz = nullptr, *gw = nullptr, *za = nullptr;
  int* body_of = nullptr;
};

static void free_mv(eb_matvec M) {
  if (!M) return;
  if (M->fmm) eb_fmm_plan_destroy(M->fmm);
  void* ptrs[] = {M->x, M->y, M->dsdth, M->rkappa, M->h, M->pk, M->dz, M->zk, M->wdz, M->gw, M->za, M->body_of};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  delete M;
}

extern "C" int eb_matvec_create(eb_matvec* out, eb_bodies B, const double* x, const double* y, const double* dz,
                                const double* dsdth, const double* rkappa, const double* zk, int32_t levmax,
                                int32_t nterms, int32_t mode, void* stream) {
  if (!out || !B || !x || !y || !dz || !dsdth || !rkappa || !zk)
    return (eb_set_error("eb_matvec_create: bad arguments"), EB_ERR_ARG);
  *out = nullptr;
  hipStream_t st = (hipStream_t)stream;
  eb_matvec M = new eb_matvec_s();
  M->B = B;
  const int n = B->n, nb = B->nbody;
  M->n = n;
  int rc;
  if ((rc = dalloc(&M->x, n)) || (rc = dalloc(&M->y, n)) || (rc = dalloc(&M->dsdth, n)) ||
      (rc = dalloc(&M->rkappa, n)) || (rc = dalloc(&M->h, nb)) || (rc = dalloc(&M->pk, nb)) ||
      (rc = dalloc(&M->dz, n)) || (rc = dalloc(&M->zk, nb)) || (rc = dalloc(&M->wdz, n)) || (rc = dalloc(&M->gw, n)) ||
      (rc = dalloc(&M->za, nb)) || (rc = dalloc(&M->body_of, n))) {
    free_mv(M);
    return rc;
  }
  std::vector<double> h(nb);
  for (int k = 0; k < nb; ++k) h[k] = 2.0 * 3.14159265358979323846 / B->nd[k];  // stokesbubb.f:446-448
  bool ok = hipMemcpyAsync(M->x, x, sizeof(double) * n, hipMemcpyDeviceToDevice, st) == hipSuccess &&
            hipMemcpyAsync(M->y, y, sizeof(double) * n, hipMemcpyDeviceToDevice, st) == hipSuccess &&
            hipMemcpyAsync(M->dsdth, dsdth, sizeof(double) * n, hipMemcpyDeviceToDevice, st) == hipSuccess &&
            hipMemcpyAsync(M->rkappa, rkappa, sizeof(double) * n, hipMemcpyDeviceToDevice, st) == hipSuccess &&
            hipMemcpyAsync(M->dz, dz, sizeof(double2) * n, hipMemcpyDeviceToDevice, st) == hipSuccess &&
            hipMemcpyAsync(M->zk, zk, sizeof(double2) * nb, hipMemcpyDeviceToDevice, st) == hipSuccess &&
            hipMemcpyAsync(M->h, h.data(), sizeof(double) * nb, hipMemcpyHostToDevice, st) == hipSuccess;
  if (!ok) {
    free_mv(M);
    return (eb_set_error("eb_matvec_create: copy failed"), EB_ERR_HIP);
  }
  hipLaunchKernelGGL(k_body_of, dim3(nb), dim3(256), 0, st, B->d_off, nb, M->body_of);
  hipLaunchKernelGGL(k_wdz, dim3((n + 255) / 256), dim3(256), 0, st, n, (const double2*)M->dz, M->h, M->body_of,
                     M->wdz);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
    free_mv(M);
    return (eb_set_error("eb_matvec_create: kernel failed"), EB_ERR_HIP);
  }
  if (mode == EB_MODE_AUTO) mode = n <= EB_DIRECT_MAX ? EB_MODE_DIRECT : EB_MODE_FMM;  // as the CADAPDR drop-in
  rc = eb_fmm_plan_create(&M->fmm, n, M->x, M->y, (const double*)M->wdz, levmax, nterms, EB_VARIANT_DR, mode, stream);
  if (rc != EB_OK) {
    M->fmm = nullptr;
    free_mv(M);
    return rc;
  }
  *out = M;
  return EB_OK;
}

extern "C" int eb_matvec_rebuild(eb_matvec M, const double* x, const double* y, const double* dz, const double* dsdth,
                                 const double* rkappa, const double* zk, void* stream) {
  if (!M || !x || !y || !dz || !dsdth || !rkappa || !zk) return (eb_set_error("eb_matvec_rebuild: bad arguments"), EB_ERR_ARG);
  hipStream_t st = (hipStream_t)stream;
  const int n = M->n, nb = M->B->nbody;
  EB_HIP_CHECK(hipMemcpyAsync(M->x, x, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  EB_HIP_CHECK(hipMemcpyAsync(M->y, y, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  EB_HIP_CHECK(hipMemcpyAsync(M->dsdth, dsdth, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  EB_HIP_CHECK(hipMemcpyAsync(M->rkappa, rkappa, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  EB_HIP_CHECK(hipMemcpyAsync(M->dz, dz, sizeof(double2) * n, hipMemcpyDeviceToDevice, st));
  EB_HIP_CHECK(hipMemcpyAsync(M->zk, zk, sizeof(double2) * nb, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_wdz, dim3((n + 255) / 256), dim3(256), 0, st, n, (const double2*)M->dz, M->h, M->body_of,
                     M->wdz);
  EB_HIP_CHECK(hipGetLastError());
  return eb_fmm_plan_rebuild(M->fmm, M->x, M->y, (const double*)M->wdz, stream);
}

extern "C" int eb_matvec_apply(eb_matvec M, const double* xx, double* yy, double* stress, double* za, double* pk,
                               void* stream) {
  if (!M || !xx || !yy) return (eb_set_error("eb_matvec_apply: bad arguments"), EB_ERR_ARG);
  hipStream_t st = (hipStream_t)stream;
  eb_bodies B = M->B;
  const double2* dens = (const double2*)xx;  // dens(i) = dcmplx(u(2i-1), u(2i)) (stokesbubb.f:527-529)
  int rc = eb_fmm_apply(M->fmm, xx, (double*)M->gw, stream);
  if (rc != EB_OK) return rc;
  hipLaunchKernelGGL(k_fasmvp_sums, dim3(B->nbody), dim3(SP_THREADS), 0, st, view(B), M->x, M->y,
                     (const double2*)M->dz, (const double2*)M->zk, M->h, dens, M->za, M->pk);
  hipLaunchKernelGGL(k_fasmvp_point, dim3((M->n + 255) / 256), dim3(256), 0, st, view(B), B->nbody, M->n, M->x, M->y,
                     (const double2*)M->dz, M->dsdth, M->rkappa, M->h, dens, (const double2*)M->za, M->pk, M->gw,
                     (double2*)yy, M->body_of);
  if (stress) EB_HIP_CHECK(hipMemcpyAsync(stress, M->gw, sizeof(double2) * M->n, hipMemcpyDeviceToDevice, st));
  if (za) EB_HIP_CHECK(hipMemcpyAsync(za, M->za, sizeof(double2) * B->nbody, hipMemcpyDeviceToDevice, st));
  if (pk) hipLaunchKernelGGL(k_shift_pk, dim3((B->nbody + 255) / 256), dim3(256), 0, st, B->nbody, pk, M->pk);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" int eb_matvec_parts(eb_matvec M, const double* xx, double* za, double* pk, void* stream) {
  if (!M || !xx) return (eb_set_error("eb_matvec_parts: bad arguments"), EB_ERR_ARG);
  hipStream_t st = (hipStream_t)stream;
  eb_bodies B = M->B;
  hipLaunchKernelGGL(k_fasmvp_sums, dim3(B->nbody), dim3(SP_THREADS), 0, st, view(B), M->x, M->y,
                     (const double2*)M->dz, (const double2*)M->zk, M->h, (const double2*)xx, M->za, M->pk);
  if (za) EB_HIP_CHECK(hipMemcpyAsync(za, M->za, sizeof(double2) * B->nbody, hipMemcpyDeviceToDevice, st));
  if (pk) hipLaunchKernelGGL(k_shift_pk, dim3((B->nbody + 255) / 256), dim3(256), 0, st, B->nbody, pk, M->pk);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" int eb_matvec_set_guard(eb_matvec M, const int32_t* flag) {
  if (!M) return (eb_set_error("eb_matvec_set_guard: bad arguments"), EB_ERR_ARG);
  return eb_fmm_set_guard(M->fmm, flag);
}

extern "C" int eb_matvec_destroy(eb_matvec M) {
  free_mv(M);
  return EB_OK;
}

// ------------------------------------------------------------------------------------------
// many independent simulations (eb_sweep)
// ------------------------------------------------------------------------------------------

// QA, QB of every point (cadapdrop.f:218-220, DR signs) in caller order and unscaled
// coordinates: the O(N^2) sum below is invariant under CADAPDR's shift and scaling
__global__ void k_sw_charges(int n, const double2* __restrict__ dens, const double2* __restrict__ wdz,
                             double2* __restrict__ qa, double2* __restrict__ qb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 d = dens[i], w = wdz[i];
  const double c = -0.15915494309189533577;  // 1/(2 pi i) = -i/(2 pi)
  const double2 dw = cmul(d, w);
  qa[i] = make_double2(-c * dw.y, c * dw.x);
  const double2 t = cadd(cscale(cmul(cconj(d), w), -1.0), cmul(d, cconj(w)));
  qb[i] = make_double2(-c * t.y, c * t.x);
}

// GW of every point from the other points of its own simulation (SLDIRDR / SLSELFDR,
// cadapdrop2.f:564-620, 724-769).  One 128-thread block per tile = (simulation, 256
// targets), two targets per thread; the simulation's sources pass through LDS in chunks of
// 256 aligned with the tiles, so only the tile's own chunk holds the targets themselves and
// runs the masked pair.  Tiles come in order of decreasing simulation size.
constexpr int SW_T = 128, SW_TILE = 2 * SW_T;
__global__ __launch_bounds__(SW_T) void k_sw_gw(const int2* __restrict__ tiles, const int* __restrict__ soff,
                                                const double* __restrict__ x, const double* __restrict__ y,
                                                const double2* __restrict__ qa, const double2* __restrict__ qb,
                                                double sgn, const int* __restrict__ mask, double2* __restrict__ gw) {
  __shared__ double2 sz[SW_TILE], sa[SW_TILE], sq[SW_TILE];
  const int2 tl = tiles[blockIdx.x];  // (simulation, first target within it)
  if (mask && !mask[tl.x]) return;    // a system the solver no longer iterates: output unused
  const int o = soff[tl.x], n = soff[tl.x + 1] - o;
  const int it0 = tl.y + (int)threadIdx.x, it1 = it0 + SW_T;
  const bool act0 = it0 < n, act1 = it1 < n;
  const double2 z0 = act0 ? make_double2(x[o + it0], y[o + it0]) : make_double2(0, 0);
  const double2 z1 = act1 ? make_double2(x[o + it1], y[o + it1]) : z0;
  double2 a10 = make_double2(0, 0), a20 = make_double2(0, 0), a11 = make_double2(0, 0), a21 = make_double2(0, 0);
  for (int c0 = 0; c0 < n; c0 += SW_TILE) {
    __syncthreads();
    for (int f = threadIdx.x; f < SW_TILE; f += SW_T) {
      const int j = c0 + f;
      if (j < n) {
        sz[f] = make_double2(x[o + j], y[o + j]);
        sa[f] = qa[o + j];
        sq[f] = qb[o + j];
      }
    }
    __syncthreads();
    const int m = min(SW_TILE, n - c0);
    if (c0 == tl.y) {
      const int s0 = threadIdx.x, s1 = s0 + SW_T;  // the targets' own positions in this chunk
      for (int jj = 0; jj < m; ++jj) {
        p2p_pair<true, true>(z0, sz[jj], sa[jj], sq[jj], jj == s0, a10, a20);
        p2p_pair<true, true>(z1, sz[jj], sa[jj], sq[jj], jj == s1 || !act1, a11, a21);
      }
    } else {
      for (int jj = 0; jj < m; ++jj) {
        p2p_pair<false, true>(z0, sz[jj], sa[jj], sq[jj], false, a10, a20);
        p2p_pair<false, true>(z1, sz[jj], sa[jj], sq[jj], false, a11, a21);
      }
    }
  }
  // GW = phi + sgn (z conj(phi') + conj(psi)): sgn = -1 for CADAPDR, +1 for CADAPMOD
  if (act0) gw[o + it0] = make_double2(a10.x + sgn * a20.x, a10.y + sgn * a20.y);
  if (act1) gw[o + it1] = make_double2(a11.x + sgn * a21.x, a11.y + sgn * a21.y);
}

// FASMVP's pointwise part (stokesbubb.f:585-611) with pinf = -pk of the simulation's first body
__global__ void k_sw_point(int n, const double* __restrict__ x, const double* __restrict__ y,
                           const double2* __restrict__ dz, const double* __restrict__ dsdth,
                           const double* __restrict__ rkappa, const double* __restrict__ h,
                           const double2* __restrict__ dens, const double2* __restrict__ za,
                           const double* __restrict__ pk, const double2* __restrict__ gw, double2* __restrict__ w,
                           const int* __restrict__ body_of, const int* __restrict__ first) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const int k = body_of[g];
  const double pi = 3.14159265358979323846;
  const double pinf = -pk[first[k]];
  const double pkk = pk[k] + pinf;
  const double2 zu = dens[g];
  const double2 d = dz[g];
  const double s = dsdth[g];
  const double2 zn = make_double2(d.y / s, -d.x / s);  // -i dz / dsdth
  const double c0 = h[k] * s * 0.5 * rkappa[g] / pi;
  const double2 zn2 = cmul(zn, zn);
  const double2 t = cmul(make_double2(zu.x, -zu.y), zn2);
  const double2 zself = make_double2(c0 * (zu.x - t.x), c0 * (zu.y - t.y));
  const double a = -0.5 * (pinf - pkk);
  const double2 pr = make_double2(-a * y[g], a * x[g]);
  const double2 st = make_double2(gw[g].x + zself.x + pr.x, gw[g].y + zself.y + pr.y);
  const double2 zak = za[k];
  w[g] = make_double2(zu.x + zak.x + st.x, zu.y + zak.y + st.y);
}

__global__ void k_sw_shift_pk(int nbody, const int* __restrict__ first, double* __restrict__ pk_out,
                              const double* __restrict__ pk) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nbody) pk_out[k] = pk[k] - pk[first[k]];
}

struct eb_sweep_s {
  eb_bodies B = nullptr;
  int nsim = 0, n = 0, ntile = 0;
  std::vector<int> uoff;  // unknowns per simulation: 2 x points
  double *x = nullptr, *y = nullptr, *dsdth = nullptr, *rkappa = nullptr, *h = nullptr, *pk = nullptr;
  double2 *dz = nullptr, *zk = nullptr, *wdz = nullptr, *gw = nullptr, *za = nullptr, *qa = nullptr, *qb = nullptr;
  int *body_of = nullptr, *first = nullptr, *soff = nullptr;
  int2* tiles = nullptr;
  const int* mask = nullptr;  // per simulation: 0 = skip in eb_sweep_apply (set around solves)
  // EB_MODE_FMM (or AUTO with a simulation above EB_DIRECT_MAX points): the layer potentials
  // on segmented FMM plans, one segment per simulation (CADAPDR's and CADAPMOD's kernels)
  int mode = EB_MODE_AUTO, maxn = 0;
  std::vector<int> soff_h;
  hipStream_t st = nullptr;
  eb_fmm_plan fdr = nullptr, fmod = nullptr;
};

static bool sw_fmm(eb_sweep S) { return S->mode == EB_MODE_FMM || (S->mode == EB_MODE_AUTO && S->maxn > EB_DIRECT_MAX); }

// the segmented plans for the current geometry (created on first use, re-binned after)
static int sw_plans(eb_sweep S, hipStream_t st) {
  if (!sw_fmm(S)) return EB_OK;
  const int d = eb::seg_levmax(S->maxn, S->nsim);
  int rc;
  if (!S->fdr) {
    // 43 terms: TOL 1e-13, as FASMVP passes it (stokesbubb.f:564-566) and INTERFACENEW (electric.f:803)
    if ((rc = eb_fmm_plan_create_segmented(&S->fdr, S->n, S->x, S->y, (const double*)S->wdz, S->nsim,
                                           S->soff_h.data(), d, 43, EB_VARIANT_DR, st)) != EB_OK)
      return rc;
    return eb_fmm_plan_create_segmented(&S->fmod, S->n, S->x, S->y, (const double*)S->wdz, S->nsim, S->soff_h.data(),
                                        d, 43, EB_VARIANT_MOD, st);
  }
  if ((rc = eb_fmm_plan_rebuild(S->fdr, S->x, S->y, (const double*)S->wdz, st)) != EB_OK) return rc;
  return eb_fmm_plan_rebuild(S->fmod, S->x, S->y, (const double*)S->wdz, st);
}

static void free_sw(eb_sweep S) {
  if (!S) return;
  if (S->fdr) eb_fmm_plan_destroy(S->fdr);
  if (S->fmod) eb_fmm_plan_destroy(S->fmod);
  void* ptrs[] = {S->x, S->y, S->dsdth, S->rkappa, S->h, S->pk, S->dz, S->zk, S->wdz, S->gw, S->za, S->qa, S->qb,
                  S->body_of, S->first, S->soff, S->tiles};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  delete S;
}

extern "C" int eb_sweep_create(eb_sweep* out, eb_bodies B, int32_t nsim, const int32_t* sim_off, const double* x,
                               const double* y, const double* dz, const double* dsdth, const double* rkappa,
                               const double* zk, void* stream) {
  if (!out || !B || nsim < 1 || !sim_off || !x || !y || !dz || !dsdth || !rkappa || !zk)
    return (eb_set_error("eb_sweep_create: bad arguments"), EB_ERR_ARG);
  *out = nullptr;
  if (sim_off[0] != 0 || sim_off[nsim] != B->nbody)
    return (eb_set_error("eb_sweep_create: simulation offsets must run from 0 to the body count"), EB_ERR_ARG);
  for (int s = 0; s < nsim; ++s)
    if (sim_off[s + 1] <= sim_off[s]) return (eb_set_error("eb_sweep_create: simulation %d has no body", s), EB_ERR_ARG);
  hipStream_t st = (hipStream_t)stream;
  eb_sweep S = new eb_sweep_s();
  S->B = B;
  S->nsim = nsim;
  const int n = B->n, nb = B->nbody;
  S->n = n;
  // point ranges, first bodies and tiles (largest simulations first)
  std::vector<int> soff(nsim + 1), first(nb);
  S->uoff.resize(nsim + 1);
  for (int s = 0; s <= nsim; ++s) {
    soff[s] = B->off[sim_off[s]];
    S->uoff[s] = 2 * soff[s];
  }
  for (int s = 0; s < nsim; ++s)
    for (int k = sim_off[s]; k < sim_off[s + 1]; ++k) first[k] = sim_off[s];
  std::vector<int> ord(nsim);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return soff[a + 1] - soff[a] > soff[b + 1] - soff[b]; });
  std::vector<int2> tiles;
  for (int s : ord)
    for (int t0 = 0; t0 < soff[s + 1] - soff[s]; t0 += SW_TILE) tiles.push_back(make_int2(s, t0));
  S->ntile = (int)tiles.size();
  S->soff_h = soff;
  S->st = st;
  for (int s2 = 0; s2 < nsim; ++s2) S->maxn = std::max(S->maxn, soff[s2 + 1] - soff[s2]);
  std::vector<double> h(nb);
  for (int k = 0; k < nb; ++k) h[k] = 2.0 * 3.14159265358979323846 / B->nd[k];  // stokesbubb.f:446-448
  int rc;
  if ((rc = dalloc(&S->x, n)) || (rc = dalloc(&S->y, n)) || (rc = dalloc(&S->dsdth, n)) ||
      (rc = dalloc(&S->rkappa, n)) || (rc = dalloc(&S->h, nb)) || (rc = dalloc(&S->pk, nb)) ||
      (rc = dalloc(&S->dz, n)) || (rc = dalloc(&S->zk, nb)) || (rc = dalloc(&S->wdz, n)) || (rc = dalloc(&S->gw, n)) ||
      (rc = dalloc(&S->za, nb)) || (rc = dalloc(&S->qa, n)) || (rc = dalloc(&S->qb, n)) ||
      (rc = dalloc(&S->body_of, n)) || (rc = dalloc(&S->first, nb)) || (rc = dalloc(&S->soff, nsim + 1)) ||
      (rc = dalloc(&S->tiles, tiles.size()))) {
    free_sw(S);
    return rc;
  }
  const bool ok = hipMemcpyAsync(S->x, x, sizeof(double) * n, hipMemcpyDeviceToDevice, st) == hipSuccess &&
                  hipMemcpyAsync(S->y, y, sizeof(double) * n, hipMemcpyDeviceToDevice, st) == hipSuccess &&
                  hipMemcpyAsync(S->dsdth, dsdth, sizeof(double) * n, hipMemcpyDeviceToDevice, st) == hipSuccess &&
                  hipMemcpyAsync(S->rkappa, rkappa, sizeof(double) * n, hipMemcpyDeviceToDevice, st) == hipSuccess &&
                  hipMemcpyAsync(S->dz, dz, sizeof(double2) * n, hipMemcpyDeviceToDevice, st) == hipSuccess &&
                  hipMemcpyAsync(S->zk, zk, sizeof(double2) * nb, hipMemcpyDeviceToDevice, st) == hipSuccess &&
                  upload(S->h, h, st) && upload(S->first, first, st) && upload(S->soff, soff, st) &&
                  upload(S->tiles, tiles, st);
  if (!ok) {
    free_sw(S);
    return (eb_set_error("eb_sweep_create: copy failed"), EB_ERR_HIP);
  }
  hipLaunchKernelGGL(k_body_of, dim3(nb), dim3(256), 0, st, B->d_off, nb, S->body_of);
  hipLaunchKernelGGL(k_wdz, dim3((n + 255) / 256), dim3(256), 0, st, n, (const double2*)S->dz, S->h, S->body_of,
                     S->wdz);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
    free_sw(S);
    return (eb_set_error("eb_sweep_create: kernel failed"), EB_ERR_HIP);
  }
  if ((rc = sw_plans(S, st)) != EB_OK) {
    free_sw(S);
    return rc;
  }
  *out = S;
  return EB_OK;
}

extern "C" int eb_sweep_set_mode(eb_sweep S, int32_t mode) {
  if (!S || (mode != EB_MODE_AUTO && mode != EB_MODE_FMM && mode != EB_MODE_DIRECT))
    return (eb_set_error("eb_sweep_set_mode: bad arguments"), EB_ERR_ARG);
  S->mode = mode;
  return sw_plans(S, S->st);
}

// za, pk (unshifted) of the density into S->za, S->pk
static void sweep_sums(eb_sweep S, const double2* dens, hipStream_t st) {
  eb_bodies B = S->B;
  hipLaunchKernelGGL(k_fasmvp_sums, dim3(B->nbody), dim3(SP_THREADS), 0, st, view(B), S->x, S->y, S->dz, S->zk, S->h,
                     dens, S->za, S->pk);
}

extern "C" int eb_sweep_apply(eb_sweep S, const double* xx, double* yy, void* stream) {
  if (!S || !xx || !yy) return (eb_set_error("eb_sweep_apply: bad arguments"), EB_ERR_ARG);
  hipStream_t st = (hipStream_t)stream;
  const int n = S->n;
  const double2* dens = (const double2*)xx;  // dens(i) = dcmplx(u(2i-1), u(2i)) (stokesbubb.f:527-529)
  if (sw_fmm(S)) {
    // every simulation's GW (the mask of converged simulations is not consulted)
    if (int rc = eb_fmm_apply(S->fdr, xx, (double*)S->gw, st)) return rc;
  } else {
    hipLaunchKernelGGL(k_sw_charges, dim3((n + 255) / 256), dim3(256), 0, st, n, dens, S->wdz, S->qa, S->qb);
    hipLaunchKernelGGL(k_sw_gw, dim3(S->ntile), dim3(SW_T), 0, st, S->tiles, S->soff, S->x, S->y, S->qa, S->qb, -1.0,
                       S->mask, S->gw);
  }
  sweep_sums(S, dens, st);
  hipLaunchKernelGGL(k_sw_point, dim3((n + 255) / 256), dim3(256), 0, st, n, S->x, S->y, S->dz, S->dsdth, S->rkappa,
                     S->h, dens, S->za, S->pk, S->gw, (double2*)yy, S->body_of, S->first);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" int eb_sweep_layer(eb_sweep S, const double* dens, double* gw, int32_t variant, void* stream) {
  if (!S || !dens || !gw || (variant != EB_VARIANT_DR && variant != EB_VARIANT_MOD))
    return (eb_set_error("eb_sweep_layer: bad arguments"), EB_ERR_ARG);
  hipStream_t st = (hipStream_t)stream;
  const int n = S->n;
  if (sw_fmm(S)) return eb_fmm_apply(variant == EB_VARIANT_DR ? S->fdr : S->fmod, dens, gw, st);
  hipLaunchKernelGGL(k_sw_charges, dim3((n + 255) / 256), dim3(256), 0, st, n, (const double2*)dens, S->wdz, S->qa,
                     S->qb);
  hipLaunchKernelGGL(k_sw_gw, dim3(S->ntile), dim3(SW_T), 0, st, S->tiles, S->soff, S->x, S->y, S->qa, S->qb,
                     variant == EB_VARIANT_DR ? -1.0 : 1.0, (const int*)nullptr, (double2*)gw);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" int eb_sweep_rebuild(eb_sweep S, const double* x, const double* y, const double* dz, const double* dsdth,
                                const double* rkappa, const double* zk, void* stream) {
  if (!S || !x || !y || !dz || !dsdth || !rkappa || !zk) return (eb_set_error("eb_sweep_rebuild: bad arguments"), EB_ERR_ARG);
  hipStream_t st = (hipStream_t)stream;
  const int n = S->n, nb = S->B->nbody;
  EB_HIP_CHECK(hipMemcpyAsync(S->x, x, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  EB_HIP_CHECK(hipMemcpyAsync(S->y, y, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  EB_HIP_CHECK(hipMemcpyAsync(S->dsdth, dsdth, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  EB_HIP_CHECK(hipMemcpyAsync(S->rkappa, rkappa, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  EB_HIP_CHECK(hipMemcpyAsync(S->dz, dz, sizeof(double2) * n, hipMemcpyDeviceToDevice, st));
  EB_HIP_CHECK(hipMemcpyAsync(S->zk, zk, sizeof(double2) * nb, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_wdz, dim3((n + 255) / 256), dim3(256), 0, st, n, (const double2*)S->dz, S->h, S->body_of,
                     S->wdz);
  EB_HIP_CHECK(hipGetLastError());
  return sw_plans(S, st);
}

extern "C" int eb_sweep_parts(eb_sweep S, const double* xx, double* za, double* pk, void* stream) {
  if (!S || !xx) return (eb_set_error("eb_sweep_parts: bad arguments"), EB_ERR_ARG);
  hipStream_t st = (hipStream_t)stream;
  sweep_sums(S, (const double2*)xx, st);
  const int nb = S->B->nbody;
  if (za) EB_HIP_CHECK(hipMemcpyAsync(za, S->za, sizeof(double2) * nb, hipMemcpyDeviceToDevice, st));
  if (pk) hipLaunchKernelGGL(k_sw_shift_pk, dim3((nb + 255) / 256), dim3(256), 0, st, nb, S->first, pk, S->pk);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" const int32_t* eb_sweep_unknown_offsets(eb_sweep S) { return S ? S->uoff.data() : nullptr; }

extern "C" int eb_sweep_set_mask(eb_sweep S, const int32_t* mask) {
  if (!S) return (eb_set_error("eb_sweep_set_mask: bad arguments"), EB_ERR_ARG);
  S->mask = mask;
  return EB_OK;
}

extern "C" int eb_sweep_destroy(eb_sweep S) {
  free_sw(S);
  return EB_OK;
}

// ------------------------------------------------------------------------------------------
// FASMVP drop-in on host arrays.  The geometry is the same for every matvec of one GMRES
// solve (MATVEC passes the /inteqn/ arrays unchanged, stokesbubb.f:449-455), so the bodies,
// the FMM plan and the device copies are kept while (ND, X, Y, DZ, DSDTH, RKAPPA, ZK, H)
// do not change.
// ------------------------------------------------------------------------------------------
namespace {
struct FasCtx {
  std::mutex mu;
  hipStream_t st = nullptr;
  eb_bodies B = nullptr;
  eb_matvec M = nullptr;
  std::vector<int> nd;
  std::vector<double> key;
  int mode = -1;  // the drop-in mode the cached matvec was built for
  double *dgeo = nullptr, *du = nullptr, *dw = nullptr, *dst = nullptr, *dza = nullptr, *dpk = nullptr;
};
FasCtx& fctx() {
  static FasCtx* c = new FasCtx();
  return *c;
}
}  // namespace

extern "C" int32_t eb_fasmvp(const int32_t* nd, const int32_t* nbk, const int32_t* k, const double* x, const double* y,
                             const double* dz, const double* dsdth, const double* rkappa, const double* zk,
                             const double* u, double* w, const double* h, double* za, double* pk, double* stress,
                             double* dens) {
  const int K = *k, N = *nbk;
  (void)h;  // h(k) = 2 pi / nd(k), as MATVEC sets it (stokesbubb.f:446-448)
  FasCtx& c = fctx();
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.st && hipStreamCreate(&c.st) != hipSuccess) return (eb_set_error("eb_fasmvp: no HIP stream"), EB_ERR_HIP);
  std::vector<double> key;
  key.reserve(7 * (size_t)N + 2 * K);
  key.insert(key.end(), x, x + N);
  key.insert(key.end(), y, y + N);
  key.insert(key.end(), dz, dz + 2 * N);
  key.insert(key.end(), dsdth, dsdth + N);
  key.insert(key.end(), rkappa, rkappa + N);
  key.insert(key.end(), zk, zk + 2 * K);
  std::vector<int> ndv(nd, nd + K);
  int rc;
  const int mode = eb_get_mode();
  if (!c.M || ndv != c.nd || key != c.key || mode != c.mode) {
    if (c.M) eb_matvec_destroy(c.M);
    if (c.B) eb_bodies_destroy(c.B);
    c.M = nullptr;
    c.B = nullptr;
    void* ptrs[] = {c.dgeo, c.du, c.dw, c.dst, c.dza, c.dpk};
    for (void* q : ptrs)
      if (q) (void)hipFree(q);
    if ((rc = eb_bodies_create(&c.B, K, nd, c.st)) != EB_OK) return rc;
    if ((rc = dalloc(&c.dgeo, key.size())) || (rc = dalloc(&c.du, 2 * (size_t)N)) || (rc = dalloc(&c.dw, 2 * (size_t)N)) ||
        (rc = dalloc(&c.dst, 2 * (size_t)N)) || (rc = dalloc(&c.dza, 2 * (size_t)K)) || (rc = dalloc(&c.dpk, K)))
      return rc;
    if (hipMemcpyAsync(c.dgeo, key.data(), sizeof(double) * key.size(), hipMemcpyHostToDevice, c.st) != hipSuccess)
      return (eb_set_error("eb_fasmvp: copy failed"), EB_ERR_HIP);
    double* g = c.dgeo;
    // the drop-in layer's mode (eb_set_mode): EB_MODE_AUTO is resolved by eb_matvec_create
    rc = eb_matvec_create(&c.M, c.B, g, g + N, g + 2 * N, g + 4 * N, g + 5 * N, g + 6 * N, 7, 43, mode, c.st);
    if (rc != EB_OK) {
      c.M = nullptr;
      return rc;
    }
    c.nd = ndv;
    c.key.swap(key);
    c.mode = mode;
  }
  if (hipMemcpyAsync(c.du, u, sizeof(double) * 2 * N, hipMemcpyHostToDevice, c.st) != hipSuccess)
    return (eb_set_error("eb_fasmvp: copy failed"), EB_ERR_HIP);
  if ((rc = eb_matvec_apply(c.M, c.du, c.dw, c.dst, c.dza, c.dpk, c.st)) != EB_OK) return rc;
  bool ok = hipMemcpyAsync(w, c.dw, sizeof(double) * 2 * N, hipMemcpyDeviceToHost, c.st) == hipSuccess &&
            (!stress || hipMemcpyAsync(stress, c.dst, sizeof(double) * 2 * N, hipMemcpyDeviceToHost, c.st) == hipSuccess) &&
            (!za || hipMemcpyAsync(za, c.dza, sizeof(double) * 2 * K, hipMemcpyDeviceToHost, c.st) == hipSuccess) &&
            (!pk || hipMemcpyAsync(pk, c.dpk, sizeof(double) * K, hipMemcpyDeviceToHost, c.st) == hipSuccess) &&
            hipStreamSynchronize(c.st) == hipSuccess;
  if (!ok) return (eb_set_error("eb_fasmvp: copy back failed"), EB_ERR_HIP);
  if (dens) memcpy(dens, u, sizeof(double) * 2 * N);
  return EB_OK;
}
