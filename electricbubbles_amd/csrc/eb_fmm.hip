// eb_fmm.hip -- MI355X (gfx950) complex-variable FMM for the Stokes layer potential of the
// reference's CADAPDR family (cadapdrop.f, cadapmod.f, cadap.f; kernels in cadapdrop2.f and
// cadap2.f), evaluated at the sources (IFLAG1 = 0) in free space (IFLAG2 = 0).
//
// Geometry, boxes and lists are the reference's: points are scaled into [0,64]^2 exactly as
// cadapdrop.f:145-205 does, level l is the uniform 2^l x 2^l grid of ASSIGN1
// (cadap2.f:292-322), the far field of a box at level l comes from the MKINTL interaction
// list (cadap2.f:780-828) and the near field from the MKNBOR 3x3 neighbourhood at the
// finest level (cadap2.f:705-742, cadapdrop2.f:623-722).  The reference re-forms every
// multipole from the sources at every level and evaluates every local at every level; here
// the same per-level expansions are produced hierarchically (P2M at the finest level, exact
// M2M/L2L translations between levels, L2P once), which reproduces the reference up to
// rounding and truncation.
//
// Expansions are stored scaled by the box half-width r of their level
// (alpha_k = A_k / r^k, lambda_m = L_m r^m), which makes the M2L operator depend on the
// box offset only, not on the level.  M2L is a real Pascal-matrix product between two
// diagonal complex scalings (the factorisation of SLSHIFT, cadap2.f:1175-1213).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <array>
#include <mutex>
#include <vector>

#include "eb_internal.h"
#include "../../include/eb_afmm.h"
#include "../../include/eb_atree.h"

using namespace eb;

// ------------------------------------------------------------------------------------------
// error reporting
// ------------------------------------------------------------------------------------------
static thread_local char g_err[512];
extern "C" void eb_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
extern "C" const char* eb_last_error(void) { return g_err; }
extern "C" int eb_version(void) { return 1; }

#define EB_MAXLEV 10
#define EB_MAXKMAX 64

// ------------------------------------------------------------------------------------------
// device helpers
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ int pidx(int ix, int iy, int W) { return (iy + 3) * W + (ix + 3); }

// [start, end) of the sorted points inside box (ix, iy) of level l, from the leaf table.
__device__ __forceinline__ void box_range(const int* __restrict__ leaf_start, int L, int l, int ix, int iy,
                                          int& s, int& e) {
  const uint32_t m = morton((uint32_t)ix, (uint32_t)iy);
  const int sh = 2 * (L - l);
  s = leaf_start[m << sh];
  e = leaf_start[(m + 1) << sh];
}

struct Geom {
  double rscal, rminx, rminy, pad;
};

// ------------------------------------------------------------------------------------------
// tree build kernels
// ------------------------------------------------------------------------------------------
constexpr int BB_THREADS = 256;

// RSCAL = 64 / max(RMAXX-RMINX, RMAXY-RMINY)  (cadapdrop.f:165-171) from np partial min/max:
// min/max are exact, so the reduction order does not change the result.
__device__ __forceinline__ void bbox_reduce(const double4* __restrict__ part, int np, Geom* __restrict__ g,
                                            double4* s) {
  double4 a = part[0];
  for (int i = threadIdx.x; i < np; i += BB_THREADS) {
    const double4 b = part[i];
    a = make_double4(fmin(a.x, b.x), fmax(a.y, b.y), fmin(a.z, b.z), fmax(a.w, b.w));
  }
  s[threadIdx.x] = a;
  __syncthreads();
  for (int off = BB_THREADS / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      const double4 b = s[threadIdx.x + off];
      a = s[threadIdx.x];
      s[threadIdx.x] = make_double4(fmin(a.x, b.x), fmax(a.y, b.y), fmin(a.z, b.z), fmax(a.w, b.w));
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a = s[0];
    const double rsx = a.y - a.x, rsy = a.w - a.z;
    const double r = (rsx > rsy) ? rsx : rsy;
    g->rscal = 64.0 / r;
    g->rminx = a.x;
    g->rminy = a.z;
  }
}

// partial min/max of x and y (cadapdrop.f:147-158), one per block; k_bbox_final reduces them.
// (A last-block reduction behind a ticket counter measured slower, 30 us against 7 + 4 us at
// 2^20 points: every block's device-scope fence costs more than the launch it saves.)
__global__ __launch_bounds__(BB_THREADS) void k_bbox_part(const double* __restrict__ x, const double* __restrict__ y,
                                                          int n, double4* __restrict__ part) {
  double mnx = x[0], mxx = x[0], mny = y[0], mxy = y[0];
  for (int i = blockIdx.x * BB_THREADS + threadIdx.x; i < n; i += gridDim.x * BB_THREADS) {
    double a = x[i], b = y[i];
    mnx = fmin(mnx, a);
    mxx = fmax(mxx, a);
    mny = fmin(mny, b);
    mxy = fmax(mxy, b);
  }
  __shared__ double4 s[BB_THREADS];
  s[threadIdx.x] = make_double4(mnx, mxx, mny, mxy);
  __syncthreads();
  for (int off = BB_THREADS / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      double4 a = s[threadIdx.x], b = s[threadIdx.x + off];
      s[threadIdx.x] = make_double4(fmin(a.x, b.x), fmax(a.y, b.y), fmin(a.z, b.z), fmax(a.w, b.w));
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// the reduction of k_bbox_part's np partials.  One block of BB_THREADS threads.
__global__ __launch_bounds__(BB_THREADS) void k_bbox_final(const double4* __restrict__ part, int np,
                                                           Geom* __restrict__ g) {
  __shared__ double4 s[BB_THREADS];
  bbox_reduce(part, np, g, s);
}

// Scaled coordinates (cadapdrop.f:203-205) and the finest-level Morton key of ASSIGN1's box
// (cadap2.f:312-319).  H = 64/2^L is a power of two, so X/H is exact and the integer
// truncation matches the reference bit for bit; coarser levels are key >> 2(L-l).
// Two neighbours of this step run in the same launch: every block first reduces the np <= 1024
// bounding-box partials of k_bbox_part itself (what k_bbox_final does; block 0 stores the
// result in gout for the later kernels), and it counts the low 8 bits of its RS_TILE_KEYS keys
// for the radix sort's first pass (hist[d * nblk + block]).
__global__ __launch_bounds__(1024) void k_keys(const double* __restrict__ x, const double* __restrict__ y,
                                                    int n, int L, const double4* __restrict__ part, int np,
                                                    Geom* __restrict__ gout, double2* __restrict__ zs0,
                                                    uint32_t* __restrict__ keys, int32_t* __restrict__ vals,
                                                    int* __restrict__ hist, int nblk) {
  __shared__ double gs[3];  // RSCAL, RMINX, RMINY
  __shared__ double4 wred[16];
  __shared__ int h[256];
  if (threadIdx.x < 256) h[threadIdx.x] = 0;
  {
    // one partial per thread (np <= 1024), then the waves' results
    double4 a = part[threadIdx.x < np ? threadIdx.x : 0];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      a.x = fmin(a.x, __shfl_xor(a.x, off, 64));
      a.y = fmax(a.y, __shfl_xor(a.y, off, 64));
      a.z = fmin(a.z, __shfl_xor(a.z, off, 64));
      a.w = fmax(a.w, __shfl_xor(a.w, off, 64));
    }
    if ((threadIdx.x & 63) == 0) wred[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
      a = wred[0];
      for (int w = 1; w < 16; ++w) {
        const double4 b = wred[w];
        a = make_double4(fmin(a.x, b.x), fmax(a.y, b.y), fmin(a.z, b.z), fmax(a.w, b.w));
      }
      // RSCAL = 64 / max(RMAXX-RMINX, RMAXY-RMINY)  (cadapdrop.f:165-171)
      const double rsx = a.y - a.x, rsy = a.w - a.z;
      const double r = (rsx > rsy) ? rsx : rsy;
      gs[0] = 64.0 / r;
      gs[1] = a.x;
      gs[2] = a.z;
      if (blockIdx.x == 0) {
        gout->rscal = gs[0];
        gout->rminx = a.x;
        gout->rminy = a.z;
      }
    }
  }
  __syncthreads();
  const double rs = gs[0], mx = gs[1], my = gs[2];
  const int nside = 1 << L;
  const double hh = 64.0 / nside;
#pragma unroll
  for (int r = 0; r < RS_TILE_KEYS / 1024; ++r) {
    const int i = blockIdx.x * RS_TILE_KEYS + r * 1024 + (int)threadIdx.x;
    if (i < n) {
      const double X = (x[i] - mx) * rs;
      const double Y = (y[i] - my) * rs;
      zs0[i] = make_double2(X, Y);
      int ix = (int)(X / hh), iy = (int)(Y / hh);
      ix = ix >= nside ? nside - 1 : ix;
      iy = iy >= nside ? nside - 1 : iy;
      ix = ix < 0 ? 0 : ix;
      iy = iy < 0 ? 0 : iy;
      const uint32_t k = morton((uint32_t)ix, (uint32_t)iy);
      keys[i] = k;
      vals[i] = i;
      atomicAdd(&h[k & 255u], 1);
    }
  }
  __syncthreads();
  if (threadIdx.x < 256) hist[threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// Segmented plans: each segment's bounding box (cadapdrop.f:147-158 on the segment's points)
// and the scale that maps its larger side onto the side T of one box of level segk.  One block
// per segment; an empty or single-point segment gets scale 1.
__global__ __launch_bounds__(BB_THREADS) void k_seg_bbox(int nseg, const int* __restrict__ seg_off,
                                                         const double* __restrict__ x, const double* __restrict__ y,
                                                         double T, double4* __restrict__ seg_geo,
                                                         double* __restrict__ seg_scale, Geom* __restrict__ g) {
  const int sg = blockIdx.x;
  const int a = seg_off[sg], b = seg_off[sg + 1];
  double mnx = INFINITY, mxx = -INFINITY, mny = INFINITY, mxy = -INFINITY;
  for (int i = a + (int)threadIdx.x; i < b; i += BB_THREADS) {
    mnx = fmin(mnx, x[i]);
    mxx = fmax(mxx, x[i]);
    mny = fmin(mny, y[i]);
    mxy = fmax(mxy, y[i]);
  }
  __shared__ double4 s[BB_THREADS];
  s[threadIdx.x] = make_double4(mnx, mxx, mny, mxy);
  __syncthreads();
  for (int off = BB_THREADS / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      const double4 u = s[threadIdx.x], v = s[threadIdx.x + off];
      s[threadIdx.x] = make_double4(fmin(u.x, v.x), fmax(u.y, v.y), fmin(u.z, v.z), fmax(u.w, v.w));
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double4 u = s[0];
    const double r = fmax(u.y - u.x, u.w - u.z);
    const double sc = (b - a >= 2 && r > 0.0) ? T / r : 1.0;
    seg_geo[sg] = make_double4(b > a ? u.x : 0.0, b > a ? u.z : 0.0, sc, 0.0);
    seg_scale[sg] = sc;
    if (sg == 0) {  // the plan-wide facts of eb_fmm_plan_info: no single scaling
      g->rscal = 1.0;
      g->rminx = 0.0;
      g->rminy = 0.0;
    }
  }
}

// Segmented plans: scaled coordinates inside the segment's box of level segk and the
// finest-level key, the segment's own leaf index clamped to its box as ASSIGN1 clamps to the
// grid (cadap2.f:312-319).
__global__ void k_seg_keys(int n, int nseg, const int* __restrict__ seg_off, const double* __restrict__ x,
                           const double* __restrict__ y, int L, int segk, const double4* __restrict__ seg_geo,
                           double2* __restrict__ zs0, uint32_t* __restrict__ keys, int32_t* __restrict__ vals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = 0, hi = nseg - 1;  // the segment of point i: the last s with seg_off[s] <= i
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (seg_off[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  const int sg = lo;
  const double4 q = seg_geo[sg];
  const double T = 64.0 / (1 << segk);
  const int sub = 1 << (L - segk);
  const double h = T / sub;
  const double X = (x[i] - q.x) * q.z, Y = (y[i] - q.y) * q.z;
  int jx = (int)(X / h), jy = (int)(Y / h);
  jx = jx >= sub ? sub - 1 : (jx < 0 ? 0 : jx);
  jy = jy >= sub ? sub - 1 : (jy < 0 ? 0 : jy);
  const int tx = (int)compact2((uint32_t)sg), ty = (int)compact2((uint32_t)sg >> 1);
  zs0[i] = make_double2(tx * T + X, ty * T + Y);
  keys[i] = morton((uint32_t)(tx * sub + jx), (uint32_t)(ty * sub + jy));
  vals[i] = i;
}

// Leaf ranges and the gather in one launch: thread t finds leaf t's first sorted point
// (leaf_start[t] = the first position whose key is >= t, t = 0..4^L; t <= nleaf) and gathers
// point t into sorted order (t < n)
__global__ void k_leaf_gather(const uint32_t* __restrict__ keys, int n, int nleaf, int* __restrict__ leaf_start,
                              const int32_t* __restrict__ perm, const double2* __restrict__ zs0,
                              const double2* __restrict__ wdz, double2* __restrict__ zs, double2* __restrict__ wdzs) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t <= nleaf) {
    int lo = 0, hi = n;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (keys[mid] < (uint32_t)t) lo = mid + 1;
      else hi = mid;
    }
    leaf_start[t] = lo;
  }
  if (t < n) {
    const int o = perm[t];
    zs[t] = zs0[o];
    wdzs[t] = wdz[o];
  }
}

__global__ void k_zero_int(int* __restrict__ a, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) a[i] = 0;
}

// number of non-empty boxes per level (the reference's NTBOX, cadap2.f:331-336): a leaf is
// counted at level l when it holds points and is the first non-empty leaf of its ancestor.
// Counts are summed per block in LDS before one atomic add per level.
__global__ __launch_bounds__(256) void k_ntbox(const int* __restrict__ leaf_start, int L, int* __restrict__ ntbox) {
  __shared__ int cnt[EB_MAXLEV + 1];
  if (threadIdx.x <= EB_MAXLEV) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  const int nleaf = 1 << (2 * L);
  if (m < nleaf) {
    const int s = leaf_start[m];
    if (leaf_start[m + 1] != s)
      for (int l = 1; l <= L; ++l) {
        const int sh = 2 * (L - l);
        if (leaf_start[(m >> sh) << sh] == s) atomicAdd(&cnt[l], 1);
      }
  }
  __syncthreads();
  if (threadIdx.x >= 1 && threadIdx.x <= L && cnt[threadIdx.x]) atomicAdd(&ntbox[threadIdx.x], cnt[threadIdx.x]);
}

// ------------------------------------------------------------------------------------------
// apply kernels
// ------------------------------------------------------------------------------------------

// QA = DENS*DZDT*H/(2 pi i), QB = (sqb*conj(DENS)*DZDT + DENS*conj(DZDT))*H/(2 pi i)
// (cadapdrop.f:218-220, cadap.f:215), in sorted order.  QC = -conj(z) QA is formed where used.
// The Hilbert product takes QA = -DENS*WDZ (its sum runs over 1/(z_i - z_j)) and no QB.
__global__ void k_charges(int n, const int32_t* __restrict__ perm, const double2* __restrict__ dens,
                          const double2* __restrict__ wdzs, double sqb, int hilbert, double2* __restrict__ qa,
                          double2* __restrict__ qb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 d = dens[perm[i]];
  const double2 w = wdzs[i];
  const double2 dw = cmul(d, w);
  if (hilbert) {
    qa[i] = make_double2(-dw.x, -dw.y);
    qb[i] = make_double2(0.0, 0.0);
    return;
  }
  const double c = -0.15915494309189533577;  // -1/(2 pi): 1/(2 pi i) = -i/(2 pi)
  qa[i] = make_double2(-c * dw.y, c * dw.x);
  const double2 t = cadd(cscale(cmul(cconj(d), w), sqb), cmul(d, cconj(w)));
  qb[i] = make_double2(-c * t.y, c * t.x);
}

// The transposed products (eb_fmm_apply_transpose).  The forward product is real-linear in the
// density: GW_i = sum_j (c_ij sigma_j + d_ij conj(sigma_j)), and its transpose under the real
// inner product Re sum conj(a_i) b_i is mu_j = sum_i (conj(c_ij) lam_i + d_ij conj(lam_i)): the
// same point set with sources and targets swapped.  Written with the forward's charges
// (alpha = WDZ/(2 pi i), gamma = conj(WDZ)/(2 pi i), QA = alpha sigma,
// QB = sqb alpha conj(sigma) + gamma sigma), for the CADAP family this is
//   mu_j = -sgn (conj(alpha_j) V_j + sqb alpha_j phi_j + conj(gamma_j) conj(phi_j))
// where phi and V = z conj(phi') + conj(psi) are the expansions' two parts at z_j for charges
// QA = lam, QB = sgn conj(lam) (and QC = -conj(z) QA as in the forward), and for the Hilbert
// product (GW_i = sum_j WDZ_j sigma_j / (z_i - z_j))
//   mu_j = conj(WDZ_j) conj(phi_j)  with QA = conj(lam).
// So the transposed product runs the forward passes unchanged on these charges, and only the
// charges kernel and the per-target output form differ.
constexpr int OUT_FORWARD = 0, OUT_T_CADAP = 1, OUT_T_HILBERT = 2;
struct OutForm {
  double sgn, sqb;
  int kind;                // OUT_*
  const double2* w;        // transposed forms: the points' WDZ in the plan's (sorted) order
  const double* segs;      // segmented plans: each segment's scale (its RSCAL), by segment
  int segk;                // segmented plans: the segments' level (segment = leaf key >> 2(L - segk))
};

// RSCAL of a point in leaf `key` of level L: the plan's, or its segment's in a segmented plan
__device__ __forceinline__ double out_scale(const OutForm& f, const Geom* g, uint32_t key, int L) {
  return f.segs ? f.segs[key >> (2 * (L - f.segk))] : g->rscal;
}

// the output at sorted point i from its two parts (phi, V); w = WDZ of the point (transposed)
template <bool TRANS>
__device__ __forceinline__ double2 out_form(const OutForm& f, double2 phi, double2 v, double2 w) {
  if constexpr (!TRANS) {
    return make_double2(phi.x + f.sgn * v.x, phi.y + f.sgn * v.y);
  } else {
    if (f.kind == OUT_T_HILBERT) return cmul(cconj(w), cconj(phi));
    const double c = 0.15915494309189533577;           // 1/(2 pi)
    const double2 al = make_double2(c * w.y, -c * w.x);   // WDZ / (2 pi i)
    const double2 ga = make_double2(-c * w.y, -c * w.x);  // conj(WDZ) / (2 pi i)
    const double2 r = cadd(cadd(cmul(cconj(al), v), cscale(cmul(al, phi), f.sqb)), cmul(cconj(ga), cconj(phi)));
    return make_double2(-f.sgn * r.x, -f.sgn * r.y);
  }
}

// the transposed products' charges from lam (caller's order) in the sorted order
__global__ void k_charges_t(int n, const int32_t* __restrict__ perm, const double2* __restrict__ lam, int kind,
                            double sgn, double2* __restrict__ qa, double2* __restrict__ qb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 l = lam[perm ? perm[i] : i];
  if (kind == OUT_T_HILBERT) {
    qa[i] = cconj(l);
    qb[i] = make_double2(0.0, 0.0);
  } else {
    qa[i] = l;
    qb[i] = make_double2(sgn * l.x, -sgn * l.y);
  }
}

// P2M at the finest level (SLMKXP, cadap2.f:444-466), one 64-lane block per leaf box.
// A chunk of up to 16 sources writes its powers t^m (t = (z-c)/r) into LDS, four lanes per
// source each covering a quarter of the exponents; output (series, k) then forms
//   alpha_{k+1} = -(1/r) sum qa t^k,  beta_{k+1} = -(1/r) sum qb t^k + k (1/r^2) sum qc t^(k-1).
template <int KMAX>
__global__ __launch_bounds__(64) void k_p2m(int L, int p, const int* __restrict__ leaf_start,
                                            const double2* __restrict__ zs, const double2* __restrict__ qa,
                                            const double2* __restrict__ qb, double2* __restrict__ mp, int W,
                                            int NBP) {
  constexpr int CH = 16, QK = KMAX / 4;
  constexpr int NO = (2 * KMAX + 63) / 64;  // outputs per lane: (series, k) = t / KMAX, t % KMAX
  __shared__ double2 P[CH][KMAX + 1];
  __shared__ double2 AB[2][CH], E[CH];
  const int nside = 1 << L;
  const int ix = blockIdx.x % nside, iy = blockIdx.x / nside;
  const uint32_t mk = morton((uint32_t)ix, (uint32_t)iy);
  const int s = leaf_start[mk], e = leaf_start[mk + 1];
  const double h = 64.0 / nside, ir = 2.0 / h;
  const double2 c = make_double2((ix + 0.5) * h, (iy + 0.5) * h);
  const int lane = threadIdx.x;
  const int js = lane & (CH - 1), qk = lane / CH;
  double2 acc[NO];
#pragma unroll
  for (int o = 0; o < NO; ++o) acc[o] = make_double2(0, 0);
  for (int c0 = s; c0 < e; c0 += CH) {
    const int nj = min(CH, e - c0);
    if (js < nj) {
      const int j = c0 + js;
      const double2 z = zs[j];
      const double2 t = make_double2((z.x - c.x) * ir, (z.y - c.y) * ir);
      if (qk == 0) {
        const double2 q1 = qa[j], q2 = qb[j];
        AB[0][js] = make_double2(-q1.x * ir, -q1.y * ir);
        AB[1][js] = make_double2(-q2.x * ir, -q2.y * ir);
        const double2 qc = cmul(make_double2(-z.x, z.y), q1);  // -conj(z) qa
        E[js] = make_double2(qc.x * ir * ir, qc.y * ir * ir);
      }
      // lane qk (0..3) writes t^(qk + 4 m), m = 0..QK-1: t^qk from t and t^2, then steps of t^4
      const double2 t2 = cmul(t, t), t4 = cmul(t2, t2);
      double2 pw = (qk & 1) ? t : make_double2(1.0, 0.0);
      if (qk & 2) pw = cmul(pw, t2);
#pragma unroll
      for (int m = 0; m < QK; ++m) {
        P[js][qk + 4 * m] = pw;
        if (m + 1 < QK) pw = cmul(pw, t4);
      }
    }
    __syncthreads();
    // every lane runs the same operations: series 0 takes its factor 0 on the QC term, so
    // the two halves of the wave do not diverge
#pragma unroll
    for (int o = 0; o < NO; ++o) {
      const int t = lane + 64 * o;
      if (t < 2 * KMAX) {
        const int sr = t / KMAX, k = t % KMAX, km1 = k > 0 ? k - 1 : 0;
        const double fk = sr ? (double)k : 0.0;
        for (int jj = 0; jj < nj; ++jj) {
          cfma(acc[o], AB[sr][jj], P[jj][k]);
          rfma(acc[o], fk, cmul(E[jj], P[jj][km1]));
        }
      }
    }
    __syncthreads();
  }
  const int pb = pidx(ix, iy, W);
#pragma unroll
  for (int o = 0; o < NO; ++o) {
    const int t = lane + 64 * o;
    if (t < 2 * KMAX) mp[(size_t)pb * 2 * KMAX + t] = (t % KMAX) < p ? acc[o] : make_double2(0, 0);
  }
}

// value of lane l - 1 (DPP wave_shr:1); lane 0 reads 0
__device__ __forceinline__ double2 from_lane_below(double2 v) {
  const int a = __builtin_amdgcn_update_dpp(0, __double2loint(v.x), 0x138, 0xF, 0xF, false);
  const int b = __builtin_amdgcn_update_dpp(0, __double2hiint(v.x), 0x138, 0xF, 0xF, false);
  const int c = __builtin_amdgcn_update_dpp(0, __double2loint(v.y), 0x138, 0xF, 0xF, false);
  const int d = __builtin_amdgcn_update_dpp(0, __double2hiint(v.y), 0x138, 0xF, 0xF, false);
  return make_double2(__hiloint2double(b, a), __hiloint2double(d, c));
}

// P2M for KMAX = 48 and 64: the same sums with lane k forming both series' coefficient k, so
// that one LDS read of t^k serves two outputs.  The QC term's sum over sources of
// E_j t_j^(k-1) is lane k - 1's sum of E_j t_j^k, taken through DPP once after the last
// source and multiplied by k then: three complex FMAs per source and lane.  Lanes k >= KMAX
// (KMAX = 48) follow along without writing.
// The sums of one box (centre c, 1/half-width ir, sorted sources s..e) into box[0..2 KMAX).
// Any KMAX that is a multiple of 8 up to 64 works; below 48 the upper lanes idle.
template <int KMAX>
__device__ __forceinline__ void p2m_box(int p, int s, int e, double2 c, double ir, const double2* __restrict__ zs,
                                        const double2* __restrict__ qa, const double2* __restrict__ qb,
                                        double2* __restrict__ box) {
  static_assert(KMAX % 8 == 0 && KMAX <= 64, "one lane per exponent");
  // chunks of 8 sources, eight lanes per source filling the powers: 7 KB of LDS per block,
  // so that six waves per SIMD fit (16-source chunks held them to three)
  constexpr int CH = 8, QK = KMAX / 8;
  __shared__ double2 P[CH][KMAX + 1];
  __shared__ double2 AB[CH][3];  // -(1/r) qa, -(1/r) qb, -(1/r^2) conj(z) qa of each source
  const int lane = threadIdx.x;
  const int js = lane & (CH - 1), qk = lane / CH;
  // lanes at or past p read row p - 1 (one broadcast address): the loop is bound by LDS
  // bandwidth, 1 KB per power read of a wave with 64 distinct rows
  const int k = lane < p ? lane : p - 1;
  const int QP = (p + 7) / 8;  // power steps per filling lane: rows up to 8 QP - 1 >= p - 1
  const double fk = (double)lane;
  double2 a0 = make_double2(0, 0), a1 = make_double2(0, 0), a2 = make_double2(0, 0);
  for (int c0 = s; c0 < e; c0 += CH) {
    const int nj = min(CH, e - c0);
    if (js < nj) {
      const int j = c0 + js;
      const double2 z = zs[j];
      const double2 t = make_double2((z.x - c.x) * ir, (z.y - c.y) * ir);
      if (qk == 0) {
        const double2 q1 = qa[j], q2 = qb[j];
        AB[js][0] = make_double2(-q1.x * ir, -q1.y * ir);
        AB[js][1] = make_double2(-q2.x * ir, -q2.y * ir);
        const double2 qc = cmul(make_double2(-z.x, z.y), q1);  // -conj(z) qa
        AB[js][2] = make_double2(qc.x * ir * ir, qc.y * ir * ir);
      }
      // lane qk (0..7) writes t^(qk + 8 m), m = 0..QK-1: t^qk from t, t^2, t^4, then steps of t^8
      const double2 t2 = cmul(t, t), t4 = cmul(t2, t2), t8 = cmul(t4, t4);
      double2 pw = (qk & 1) ? t : make_double2(1.0, 0.0);
      if (qk & 2) pw = cmul(pw, t2);
      if (qk & 4) pw = cmul(pw, t4);
#pragma unroll
      for (int m = 0; m < QK; ++m) {
        if (m < QP) P[js][qk + 8 * m] = pw;
        if (m + 1 < QK && m + 1 < QP) pw = cmul(pw, t8);
      }
    }
    __syncthreads();
    for (int jj = 0; jj < nj; ++jj) {
      const double2 pk = P[jj][k];
      cfma(a0, AB[jj][0], pk);
      cfma(a1, AB[jj][1], pk);
      cfma(a2, AB[jj][2], pk);
    }
    __syncthreads();
  }
  rfma(a1, fk, from_lane_below(a2));  // lane 0 takes 0 (and fk = 0)
  if (lane < KMAX) {
    const bool in = lane < p;
    box[lane] = in ? a0 : make_double2(0, 0);
    box[KMAX + lane] = in ? a1 : make_double2(0, 0);
  }
}

template <int KMAX>
__global__ __launch_bounds__(64) void k_p2m_wide(int L, int p, const int* __restrict__ leaf_start,
                                                 const double2* __restrict__ zs, const double2* __restrict__ qa,
                                                 const double2* __restrict__ qb, double2* __restrict__ mp, int W,
                                                 int NBP) {
  static_assert(KMAX >= 48 && KMAX <= 64, "one lane per exponent");
  const int nside = 1 << L;
  const int ix = blockIdx.x % nside, iy = blockIdx.x / nside;
  const uint32_t mk = morton((uint32_t)ix, (uint32_t)iy);
  const int s = leaf_start[mk], e = leaf_start[mk + 1];
  const double h = 64.0 / nside, ir = 2.0 / h;
  const double2 c = make_double2((ix + 0.5) * h, (iy + 0.5) * h);
  p2m_box<KMAX>(p, s, e, c, ir, zs, qa, qb, mp + (size_t)pidx(ix, iy, W) * 2 * KMAX);
}

// ------------------------------------------------------------------------------------------
// Compacted M2L target lists.  For parity class c: the non-empty boxes of levels
// 2..L, entry ((l - 2) << 24) | q with q the box's index within the class at level l, in level
// then q order; each level's run is padded with -1 to a multiple of 4 so that every group of
// four entries (one wave of k_xlat) lies within one level.  count[c] = the padded length.
// The list of class c starts at c * cap, cap = sum over levels of (boxes of the class + 3).
// ------------------------------------------------------------------------------------------
// Two passes over blocks of 1024 boxes, block blk = cls * nb + j, j enumerating the chunks of
// levels 2..L in order: m2l_count_block counts each chunk's non-empty boxes, m2l_write_block
// places them after the chunks (and padded levels) before it.  Both run as the first 4 * nb
// blocks of k_tree_count and k_tree_write.
__device__ __forceinline__ void m2l_chunk(int blk, int L, int nb, int& cls, int& l, int& c0) {
  cls = blk / nb;
  int j = blk - cls * nb;
  for (l = 2; l <= L; ++l) {
    const int nc = ((1 << (2 * (l - 1))) + 1023) >> 10;
    if (j < nc) break;
    j -= nc;
  }
  c0 = j << 10;
}

__device__ __forceinline__ bool m2l_live(const int* __restrict__ leaf_start, int L, int l, int cls, int q) {
  const int half = 1 << (l - 1);
  if (q >= half * half) return false;
  const int ix = 2 * (q % half) + (cls & 1), iy = 2 * (q / half) + (cls >> 1);
  int bs, be;
  box_range(leaf_start, L, l, ix, iy, bs, be);
  return be > bs;
}

__device__ __forceinline__ void m2l_count_block(int blk, int L, int nb, int lmin, const int* __restrict__ leaf_start,
                                                int* __restrict__ ccnt) {
  __shared__ int wsum[16];
  int cls, l, c0;
  m2l_chunk(blk, L, nb, cls, l, c0);
  const bool live = l >= lmin && m2l_live(leaf_start, L, l, cls, c0 + (int)threadIdx.x);
  const unsigned long long m = __ballot(live);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int k = 0; k < 16; ++k) acc += wsum[k];
    ccnt[blk] = acc;
  }
}

__device__ __forceinline__ void m2l_write_block(int blk, int L, int nb, int lmin, const int* __restrict__ leaf_start,
                                                const int* __restrict__ ccnt, int cap, int* __restrict__ list,
                                                int* __restrict__ count) {
  __shared__ int wsum[16], wpre[16], base_s, last_s, lvl_s;
  int cls, l, c0;
  m2l_chunk(blk, L, nb, cls, l, c0);
  if (threadIdx.x == 0) {
    // entries before this chunk: whole padded levels below l, then this level's earlier chunks
    const int* cc = ccnt + cls * nb;
    int pos = 0, j = 0;
    for (int ll = 2; ll < l; ++ll) {
      const int nc = ((1 << (2 * (ll - 1))) + 1023) >> 10;
      int t = 0;
      for (int k = 0; k < nc; ++k) t += cc[j + k];
      pos += (t + 3) & ~3;
      j += nc;
    }
    const int nc = ((1 << (2 * (l - 1))) + 1023) >> 10;
    const int mine = c0 >> 10;
    int lvl = 0;
    for (int k = 0; k < nc; ++k) {
      if (k == mine) base_s = pos + lvl;
      lvl += cc[j + k];
    }
    last_s = mine == nc - 1;
    lvl_s = pos + lvl;  // end of this level's entries before padding
  }
  const int q = c0 + (int)threadIdx.x;
  const bool live = l >= lmin && m2l_live(leaf_start, L, l, cls, q);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const unsigned long long m = __ballot(live);
  if (lane == 0) wsum[wv] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int k = 0; k < 16; ++k) {
      wpre[k] = acc;
      acc += wsum[k];
    }
  }
  __syncthreads();
  int* out = list + (size_t)cls * cap;
  if (live) out[base_s + wpre[wv] + __popcll(m & ((1ull << lane) - 1ull))] = ((l - 2) << 24) | q;
  if (last_s) {
    const int end = lvl_s, pad = (4 - ((end) & 3)) & 3;
    if ((int)threadIdx.x < pad) out[end + threadIdx.x] = -1;
    if (l == L && threadIdx.x == 0) count[cls] = end + pad;
  }
}

// ------------------------------------------------------------------------------------------
// Translations on FP64 matrix cores: M2M (upward), M2L (interaction lists), L2L (downward).
//
// With expansions scaled by the half-width of their box, each translation is
//   out_m = post_m * sum_k Mat[k][m] * (pre_k * in_k)
// with complex diagonal pre/post scalings and a real Pascal matrix Mat:
//   M2L  (SLSHIFT, cadap2.f:1175-1213): Mat[k][m] = C(m+k,k), pre_k = u^(k+1), post_m = (-u)^m,
//        u = r/(c_target - c_source), 27 sources per target (MKINTL, cadap2.f:780-828);
//   M2M: Mat[k][n] = C(n,k), pre_k = 2^-(k+1) d^-k, post_n = d^n, 4 children per parent;
//   L2L: Mat[m][j] = C(m,j), pre_m = d^m, post_j = (2d)^-j, 1 parent per child;
// d = (child centre - parent centre)/r_parent = (+-1 +- i)/2.
// The real product runs on v_mfma_f64_16x16x4f64 with A = Mat (LDS) and B = the prescaled
// input as real columns; the scalings run on the VALU.  A wave owns 4 target boxes of one
// parity class (so the source offsets and scalings are wave-uniform); its 16 B columns are
// (target, series, re/im).  Operand layout (lane l, register r):
//   A[l%16][l/16],  B[l/16][l%16],  D[l/16 + 4r][l%16].
// ------------------------------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));

enum { XL_M2L = 0, XL_M2M = 1, XL_L2L = 2 };

struct XlatLevels {
  int nent;                   // target levels handled by the launch
  int L;
  int total;                  // logical blocks
  int lev[EB_MAXLEV];
  int blk0[EB_MAXLEV + 1];    // first logical block of each entry
  int Wt[EB_MAXLEV], NBPt[EB_MAXLEV];
  int Ws[EB_MAXLEV], NBPs[EB_MAXLEV];
  long long offt[EB_MAXLEV], offs[EB_MAXLEV];  // expansion offsets (double2 elements)
  int segk;                   // segmented plans: the segments' level (M2L sources outside the
                              // target's segment read as zero); 0 otherwise
};

// value of lane l^1 (DPP quad_perm [1,0,3,2])
__device__ __forceinline__ double swap_adjacent(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_mov_dpp(lo, 0xB1, 0xF, 0xF, false);
  hi = __builtin_amdgcn_mov_dpp(hi, 0xB1, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// The per-wave part of a translation: this wave's four targets (target col / 4 of lane col), box
// pt of the destination level (dst), each from the NS sources that src_box(o) names in the
// source level (src, nsrc_boxes boxes; an index past it reads zeros); live lanes write.
// Cs, Ps and Qs are the LDS tables (Ps and Qs at this wave's rows).  No barrier inside: a
// kernel may call it for several target sets between its own barriers.
// MTW < MT: the wave computes only output tiles mt0 .. mt0 + MTW - 1 (and the tail rows when
// tail is set), so that several waves can share one target set's rows.  JIT: read the
// post-scaling weights where they are used instead of in one batch per source (below).
template <int MODE, int KMAX, int KS, int MT, int TR, int MTW = MT, bool JIT = false, typename SrcBox>
__device__ __forceinline__ void xlat_wave(const double* Cs, const double2* Ps, const double2* Qs, int pt, bool live,
                                          SrcBox src_box, const double2* __restrict__ src, int nsrc_boxes,
                                          double2* __restrict__ dst, int mt0 = 0, bool tail = true) {
  constexpr int NS = MODE == XL_M2L ? 27 : (MODE == XL_M2M ? 4 : 1);
  constexpr int NM = MT * 16 + TR, NK = KS * 4;
  constexpr int CS = TR ? (NM + 16) / 32 * 32 + 16 : NM + 4;
  const int lane = threadIdx.x & 63;
  const int col = lane & 15, kr = lane >> 4;
  const int s = (col >> 1) & 1, comp = col & 1;
  // Source expansions through a buffer resource covering both series of the level: the
  // per-k row step is a scalar offset (no per-lane address arithmetic per load), and an
  // offset outside the level reads 0 instead of faulting.
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
      (void*)src, (short)0, (int)(uint32_t)((size_t)2 * KMAX * nsrc_boxes * 16), 0x00020000);
  const int row_bytes = 4 * 16;  // k -> k + 4 within the box's coefficients
  const int lane_row = (s * KMAX + kr) * 16;
  // the first n k steps of source o
  auto load_src = [&](int o, double2* a, auto n) {
    constexpr int N = decltype(n)::value;
    const int off = lane_row + src_box(o) * (2 * KMAX * 16);
#pragma unroll
    for (int ks = 0; ks < N; ++ks)
      a[ks] = __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(rsrc, off, ks * row_bytes, 0));
  };
  // M2L: the sources after the first M2L_NEAR (m2l_offsets: convergence ratio at most 1/2,
  // against sqrt(2)/2 for the first group) take only the first KSF = ceil(2 KS / 3) k steps
  // of their expansions, with every output row kept: 24 of the 36 terms at p = 33..36, where
  // the 2^20-point circles stay at 8.2e-12 against 7.1e-12 with all terms (20 terms: 4e-10;
  // keeping 20 output rows instead, whatever the input: 4e-10).  The M2L's matrix-core work
  // drops by 28 %: 459 -> 380 us at 2^20 points, p = 36.
  constexpr bool SPLIT = MODE == XL_M2L;
  constexpr int NNEAR = SPLIT ? M2L_NEAR : NS;
  constexpr int KSF = SPLIT ? (2 * KS + 2) / 3 : KS;
  // out_m = Re/Im(w_m y_m) with y_m split over this lane (v) and its partner (v'): summed
  // over the sources it is A -+ B' with A = sum w.x v and B = sum w.y v of the partner lane,
  // so one lane swap after the last source replaces one per source.
  double sa[MTW][4], sb[MTW][4];
#pragma unroll
  for (int mt = 0; mt < MTW; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) sa[mt][r] = sb[mt][r] = 0.0;
  // Tail rows 0..3 (TR = 4; the tiles then hold rows 4..4 + 16 MT - 1): this lane's partial
  // sums over its k rows k = 4 ks + kr, reduced over the four k-row lanes of its column after
  // the last source.  In these rows the entries are polynomials of degree r in ks for fixed
  // kr: C(r + k, k) (M2L) and C(k, r) (L2L), so sum_ks Mat[k][r] xb[ks] =
  // sum_d et[r][d] B_d with the binomial moments B_d = sum_ks C(ks, d) xb[ks] (integer
  // constants) and et[r][d] the d-th forward difference over ks of Mat[4 ks + kr][r] at 0
  // (exact integers): ten coefficients per lane instead of the 4 KS table entries.  M2M's
  // rows 0..3 have C(n, k) = 0 for k > n: only ks = 0 contributes.
  double ta[TR ? TR : 1], tb[TR ? TR : 1];
#pragma unroll
  for (int r = 0; r < TR; ++r) ta[r] = tb[r] = 0.0;
  double et[TR ? 10 : 1];
  if constexpr (TR > 0) {
    if constexpr (MODE == XL_M2M) {
#pragma unroll
      for (int r = 0; r < 4; ++r) et[r] = Cs[kr * CS + r];
    } else {
      int i = 0;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = Cs[(4 * q + kr) * CS + r];
#pragma unroll
        for (int d = 0; d <= r; ++d) {
          et[i++] = v[0];
#pragma unroll
          for (int q = 0; q + 1 < 4 - d; ++q) v[q] = v[q + 1] - v[q];
        }
      }
    }
  }
  // One source: prescale the loaded expansion (cur), start loading the next source into the
  // same registers (the prescaled copy xb is all the step needs; the loads are in flight
  // during the products: a second register set measured slower, 0.479 against 0.469 ms at
  // 2^20 points, p = 36), read all of the step's post-scaling weights at once (read one by
  // one beside their FMAs, each waited for, they cost 0.457 against 0.469 ms), the Pascal
  // product on the matrix cores, then the post-scaling.  The empty asm statement is a
  // compiler barrier for memory that keeps the Pascal tile in LDS, re-read per source
  // instead of held in 2*KS*MT registers.
  auto step = [&](int o, double2* cur, auto ksg, auto mtg, auto ksn) {
    constexpr int KG = decltype(ksg)::value, MG = decltype(mtg)::value;
    __asm__ volatile("" ::: "memory");
    double xb[KG];
#pragma unroll
    for (int ks = 0; ks < KG; ++ks) {
      const double2 w = Ps[2 * (o * NK + 4 * ks + kr) + comp];
      xb[ks] = fma(w.x, cur[ks].x, w.y * cur[ks].y);
    }
    if constexpr (decltype(ksn)::value > 0) load_src(o + 1, cur, ksn);
    // JIT (the one-level L2L kernels, uniform and adaptive): the registers the batch would hold
    // cost them occupancy (the L2L stage 44 -> 38 us at 2^20 points, adaptive 104 -> 97 us), while the M2L's batch pays (three waves per SIMD
    // for the M2L, with the weights read where used, measured 381 -> 389 us) and so does the
    // L2L subtree kernel's (its few waves' chains wait on the reads: 9 -> 12 us at p = 37..43)
    double2 qw[JIT ? 1 : MG][4];
    double2 tw[TR && !JIT ? TR : 1];
    if constexpr (!JIT) {
#pragma unroll
      for (int mt = 0; mt < MG; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) qw[mt][r] = Qs[o * NM + TR + 16 * (mt0 + mt) + kr + 4 * r];
#pragma unroll
      for (int r = 0; r < TR; ++r) tw[r] = Qs[o * NM + r];
      __builtin_amdgcn_sched_barrier(0);
    }
    v4d acc[MG];
#pragma unroll
    for (int mt = 0; mt < MG; ++mt) acc[mt] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < KG; ++ks)
#pragma unroll
      for (int mt = 0; mt < MG; ++mt)
        acc[mt] = __builtin_amdgcn_mfma_f64_16x16x4f64(Cs[(4 * ks + kr) * CS + TR + 16 * (mt0 + mt) + col], xb[ks], acc[mt],
                                                       0, 0, 0);
    if constexpr (TR > 0) {
      double part[4];
      if constexpr (MODE == XL_M2M) {
#pragma unroll
        for (int r = 0; r < 4; ++r) part[r] = et[r] * xb[0];
      } else {
        double b0 = xb[0], b1 = 0.0, b2 = 0.0, b3 = 0.0;
#pragma unroll
        for (int ks = 1; ks < KG; ++ks) {
          b0 += xb[ks];
          b1 = fma((double)ks, xb[ks], b1);
          if (ks >= 2) b2 = fma((double)(ks * (ks - 1) / 2), xb[ks], b2);
          if (ks >= 3) b3 = fma((double)(ks * (ks - 1) * (ks - 2) / 6), xb[ks], b3);
        }
        part[0] = et[0] * b0;
        part[1] = fma(et[1], b0, et[2] * b1);
        part[2] = fma(et[3], b0, fma(et[4], b1, et[5] * b2));
        part[3] = fma(et[6], b0, fma(et[7], b1, fma(et[8], b2, et[9] * b3)));
      }
#pragma unroll
      for (int r = 0; r < TR; ++r) {
        const double2 w = JIT ? Qs[o * NM + r] : tw[JIT ? 0 : r];
        ta[r] = fma(w.x, part[r], ta[r]);
        tb[r] = fma(w.y, part[r], tb[r]);
      }
    }
#pragma unroll
    for (int mt = 0; mt < MG; ++mt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double2 w = JIT ? Qs[o * NM + TR + 16 * (mt0 + mt) + kr + 4 * r] : qw[JIT ? 0 : mt][r];
        sa[mt][r] = fma(w.x, acc[mt][r], sa[mt][r]);
        sb[mt][r] = fma(w.y, acc[mt][r], sb[mt][r]);
      }
  };
  using IKS = std::integral_constant<int, KS>;
  using IMT = std::integral_constant<int, MTW>;
  using IKSF = std::integral_constant<int, KSF>;
  using INONE = std::integral_constant<int, 0>;
  double2 ra[KS];
  load_src(0, ra, IKS{});
  if constexpr (NNEAR < NS) {
#pragma unroll 1
    for (int o = 0; o < NNEAR - 1; ++o) step(o, ra, IKS{}, IMT{}, IKS{});
    step(NNEAR - 1, ra, IKS{}, IMT{}, IKSF{});
#pragma unroll 1
    for (int o = NNEAR; o < NS - 1; ++o) step(o, ra, IKSF{}, IMT{}, IKSF{});
    step(NS - 1, ra, IKSF{}, IMT{}, INONE{});
  } else {
#pragma unroll 1
    for (int o = 0; o < NS - 1; ++o) step(o, ra, IKS{}, IMT{}, IKS{});
    step(NS - 1, ra, IKS{}, IMT{}, INONE{});
  }
  const uint32_t sgn_hi = comp ? 0u : 0x80000000u;  // A - B' on component 0, A + B' on 1
  double lam[MTW][4];
#pragma unroll
  for (int mt = 0; mt < MTW; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {  // every lane takes part in the swap
      const double bo = swap_adjacent(sb[mt][r]);
      lam[mt][r] = sa[mt][r] + __hiloint2double(__double2hiint(bo) ^ (int)sgn_hi, __double2loint(bo));
    }
  // tail: sum over the k-row lanes (lanes col + 16 kr), then lane kr keeps rows kr + 4 r'
  double tlam[TR ? TR / 4 : 1];
  if constexpr (TR > 0) {
#pragma unroll
    for (int r = 0; r < TR; ++r) {
      ta[r] += __shfl_xor(ta[r], 16, 64);
      tb[r] += __shfl_xor(tb[r], 16, 64);
      ta[r] += __shfl_xor(ta[r], 32, 64);
      tb[r] += __shfl_xor(tb[r], 32, 64);
    }
#pragma unroll
    for (int q = 0; q < TR / 4; ++q) {
      double a = ta[4 * q], bq = tb[4 * q];
#pragma unroll
      for (int j = 1; j < 4; ++j) {
        a = kr == j ? ta[4 * q + j] : a;
        bq = kr == j ? tb[4 * q + j] : bq;
      }
      const double bo = swap_adjacent(bq);
      tlam[q] = a + __hiloint2double(__double2hiint(bo) ^ (int)sgn_hi, __double2loint(bo));
    }
  }
  if (live) {
    double* out = (double*)dst;
#pragma unroll
    for (int mt = 0; mt < MTW; ++mt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = TR + 16 * (mt0 + mt) + kr + 4 * r;
        const size_t at = 2 * ((size_t)pt * 2 * KMAX + s * KMAX + m) + comp;
        out[at] = MODE == XL_L2L ? out[at] + lam[mt][r] : lam[mt][r];
      }
    if (tail) {  // every wave forms the tail rows; one of several sharing a target set writes them
#pragma unroll
      for (int q = 0; q < TR / 4; ++q) {
        const int m = kr + 4 * q;
        const size_t at = 2 * ((size_t)pt * 2 * KMAX + s * KMAX + m) + comp;
        out[at] = MODE == XL_L2L ? out[at] + tlam[q] : tlam[q];
      }
    }
  }
}

// With a compacted target list (M2L only: m2l_list / m2l_count from k_tree_write) block b
// takes entries 16 (b / 4) .. 16 (b / 4) + 15 of parity class b % 4, four per wave: only
// non-empty target boxes are translated, which matters where most boxes are empty (drops
// scattered over a large square).  Blocks past the class's count leave at once.
//
// Output rows: with TR = 4, rows 0..3 on the vector ALUs (from four binomial moments of the
// lane's inputs, see xlat_wave) and MT tiles of 16 on the matrix cores for rows 4..; with
// TR = 0, the tiles alone from row 0.  A 16x16x4 FP64 MFMA costs as much as 64 vector FMAs
// per lane and issues on the same SIMD (they do not overlap); the tail costs about
// 3 KS + 18 per source, so p = 33..36 pays for 36 rows instead of 48.  Rows past 16 MT + TR
// are not written: they are zero for every p <= 16 MT + TR (the tables vanish there), and no
// reader goes past row 4 KS - 1 (translations) or p - 1 (k_l2p).
//
// ADAPT: boxes of an adaptive tree (the adaptive plan below), stored box-major by box index.
// Targets come from per-class lists (m2l_list / m2l_count / m2l_cap, no level runs: the
// scaled translations do not depend on the level), sources from a slot table stab[target * NS
// + slot]; a slot without a box holds nsrc, one past the last box, which the buffer resource
// reads as zeros.
template <int MODE, int KMAX, int KS, int MT, int TR, bool ADAPT = false, bool SEG = false>
__device__ __forceinline__ void xlat_body(const XlatLevels& LS, const int* __restrict__ leaf_start,
                                          const double2* __restrict__ src, double2* __restrict__ dst,
                                          const double* __restrict__ Mat, const double2* __restrict__ pre,
                                          const double2* __restrict__ post, const int* __restrict__ offs,
                                          const int* __restrict__ m2l_list, const int* __restrict__ m2l_count,
                                          int m2l_cap, const int* __restrict__ stab = nullptr, int nsrc = 0) {
  static_assert(KS * 4 <= KMAX && MT * 16 + TR <= KMAX, "tile exceeds the padded expansion length");
  static_assert((TR == 0 || TR == 4) && KS * 4 <= MT * 16 + TR, "tail rows must cover the expansion");
  constexpr int NS = MODE == XL_M2L ? 27 : (MODE == XL_M2M ? 4 : 1);
  constexpr int NM = MT * 16 + TR, NK = KS * 4;
  // row stride of the Pascal tile: 2 CS = 32 (mod 64) dwords puts the four k rows that one
  // MFMA operand read touches on alternate halves of the LDS banks
  constexpr int CS = TR ? (NM + 16) / 32 * 32 + 16 : NM + 4;
  __shared__ __align__(16) double Cs[NK * CS];
  __shared__ double2 Ps[NS * NK * 2];  // per (source, k, component): the weights of this lane's column
  __shared__ double2 Qs[NS * NM];
  __shared__ int Sx[ADAPT ? 16 * NS : 1];  // ADAPT: the block's 16 targets' source boxes
  // XCD-aware order: consecutive logical blocks (neighbouring rows, all four parity classes)
  // are dispatched to the same XCD and share its L2.
  const bool compact = ADAPT || (MODE == XL_M2L && m2l_list != nullptr);
  // waves per block: 4 on every path; the compacted M2L also runs with 12 (tpb = 4 wpb
  // targets of one class per block)
  const int wpb = blockDim.x >> 6, tpb = 4 * wpb;
  int nphys = gridDim.x;
  if (compact) {
    // logical blocks that can hold entries: 4 classes x the longest class's blocks, so that
    // the XCD bands below spread them over all eight XCDs
    int mx = 0;
    for (int c = 0; c < 4; ++c) mx = max(mx, (m2l_count[c] + tpb - 1) / tpb);
    nphys = (4 * mx + 7) / 8 * 8;
    if ((int)blockIdx.x >= nphys) return;  // the whole block: before any barrier
  }
  const int b = (blockIdx.x % 8) * (nphys / 8) + blockIdx.x / 8;
  const int cls = b & 3;
  const int ccount = compact ? m2l_count[cls] : 0;
  if (compact && tpb * (b >> 2) >= ccount) return;
  if constexpr (ADAPT) {
    // the slot table's rows of this block's targets, read once instead of per source step
    for (int i = threadIdx.x; i < 16 * NS; i += blockDim.x) {
      const int idx = tpb * (b >> 2) + i / NS;
      const int ent = idx < ccount ? m2l_list[(size_t)cls * m2l_cap + idx] : -1;
      Sx[i] = ent >= 0 ? stab[(size_t)ent * NS + i % NS] : nsrc;
    }
  }
  for (int i = threadIdx.x; i < NK * NM; i += blockDim.x) {
    const int k = i / NM, m = i % NM;
    Cs[k * CS + m] = Mat[k * KMAX + m];
  }
  // a lane of component c forms c ? Im(w a) : Re(w a) = W.x a.x + W.y a.y with
  // W = c ? (w.y, w.x) : (w.x, -w.y): the component choice sits in the table, not in selects
  for (int i = threadIdx.x; i < NS * NK; i += blockDim.x) {
    const int o = i / NK, k = i % NK;
    const int row = MODE == XL_M2L ? cls * 27 + o : (MODE == XL_M2M ? o : cls);
    const double2 w = pre[(size_t)row * KMAX + k];
    Ps[2 * i] = make_double2(w.x, -w.y);
    Ps[2 * i + 1] = make_double2(w.y, w.x);
  }
  for (int i = threadIdx.x; i < NS * NM; i += blockDim.x) {
    const int o = i / NM, m = i % NM;
    const int row = MODE == XL_M2L ? cls * 27 + o : (MODE == XL_M2M ? o : cls);
    Qs[i] = post[(size_t)row * KMAX + m];
  }
  __syncthreads();
  if (b >= LS.total) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = lane & 15, kr = lane >> 4;
  const int t = col >> 2, s = (col >> 1) & 1, comp = col & 1;
  int e = 0, q;
  bool valid;
  if (compact) {
    const int idx = tpb * (b >> 2) + 4 * wv + t;
    const int ent = idx < ccount ? m2l_list[(size_t)cls * m2l_cap + idx] : -1;
    valid = ent >= 0;
    if constexpr (ADAPT) {
      if (__ballot(valid) == 0) return;  // past the class's list; no barrier follows
      q = valid ? ent : 0;
    } else {
      // every valid entry of a wave has the same level (runs are padded to groups of 4)
      int ev = valid ? (ent >> 24) : -1;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) ev = max(ev, __shfl_xor(ev, o, 64));
      if (ev < 0) return;  // a wave of padding only; no barrier follows
      e = __builtin_amdgcn_readfirstlane(ev);
      q = valid ? (ent & 0xFFFFFF) : 0;
    }
  } else {
    while (e + 1 < LS.nent && b >= LS.blk0[e + 1]) ++e;
    const int g = (b - LS.blk0[e]) >> 2;
    q = (g * 4 + wv) * 4 + t;
    valid = q < (1 << (2 * (LS.lev[e] - 1)));
  }
  const int l = LS.lev[e], Wt = LS.Wt[e], NBPt = LS.NBPt[e], Ws = LS.Ws[e], NBPs = LS.NBPs[e];
  const int half = ADAPT ? 1 : 1 << (l - 1);  // the level tables are not read in ADAPT
  const int qq = valid ? q : 0;
  const int ix = 2 * (qq % half) + (cls & 1), iy = 2 * (qq / half) + (cls >> 1);
  bool live = valid;
  if (!ADAPT && MODE != XL_M2M && valid) {
    int bs, be;
    box_range(leaf_start, LS.L, l, ix, iy, bs, be);
    live = be > bs;
  }
  // A wave whose four target boxes are all empty has nothing to write (their local
  // expansions are never read: L2L and L2P only visit non-empty boxes).  No barrier follows,
  // so the whole wave may leave; sparse geometries leave most waves here.
  if (!ADAPT && MODE != XL_M2M && __ballot(live) == 0) return;
  const int pt = ADAPT ? qq : pidx(ix, iy, Wt);
  const int* sxp = &Sx[ADAPT ? (4 * wv + t) * NS : 0];
  auto src_box = [&](int o) -> int {
    if constexpr (ADAPT) return sxp[o];
    if constexpr (MODE == XL_M2L && SEG) {
      // a source in another segment's box of level segk reads as zero (an index past the level)
      const int dx = offs[(cls * 27 + o) * 2 + 0], dy = offs[(cls * 27 + o) * 2 + 1];
      const int sh = l - LS.segk;
      const bool same = (((ix + dx) >> sh) == (ix >> sh)) && (((iy + dy) >> sh) == (iy >> sh));
      return same ? pt + dy * Ws + dx : NBPs;
    }
    if (MODE == XL_M2L) return pt + offs[(cls * 27 + o) * 2 + 1] * Ws + offs[(cls * 27 + o) * 2 + 0];
    if (MODE == XL_M2M) return pidx(2 * ix + (o & 1), 2 * iy + (o >> 1), Ws);
    return pidx(ix >> 1, iy >> 1, Ws);
  };
  xlat_wave<MODE, KMAX, KS, MT, TR, MT, MODE == XL_L2L>(Cs, Ps, Qs, pt, live, src_box, ADAPT ? src : src + LS.offs[e],
                                                       ADAPT ? nsrc : NBPs,
                                    ADAPT ? dst : dst + LS.offt[e]);
}

#define EB_XLAT_ARGS                                                                                     \
  XlatLevels LS, const int* __restrict__ leaf_start, const double2* __restrict__ src,                  \
      double2* __restrict__ dst, const double* __restrict__ Mat, const double2* __restrict__ pre,      \
      const double2* __restrict__ post, const int* __restrict__ offs, const int* __restrict__ m2l_list, \
      const int* __restrict__ m2l_count, int m2l_cap
template <int MODE, int KMAX, int KS, int MT, int TR, bool SEG = false>
__global__ __launch_bounds__(256) void k_xlat(EB_XLAT_ARGS) {
  xlat_body<MODE, KMAX, KS, MT, TR, false, SEG>(LS, leaf_start, src, dst, Mat, pre, post, offs, m2l_list, m2l_count,
                                               m2l_cap);
}
// The M2L with a vector tail: its LDS admits two workgroups per CU, so the registers are held
// to what two waves per SIMD allow.
template <int MODE, int KMAX, int KS, int MT, int TR, bool SEG = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_xlat2(EB_XLAT_ARGS) {
  xlat_body<MODE, KMAX, KS, MT, TR, false, SEG>(LS, leaf_start, src, dst, Mat, pre, post, offs, m2l_list, m2l_count,
                                               m2l_cap);
}
// the same on the boxes of an adaptive tree (ADAPT above)
template <int MODE, int KMAX, int KS, int MT, int TR>
__global__ __launch_bounds__(256) void k_axlat(EB_XLAT_ARGS, const int* __restrict__ stab, int nsrc) {
  xlat_body<MODE, KMAX, KS, MT, TR, true>(LS, leaf_start, src, dst, Mat, pre, post, offs, m2l_list, m2l_count,
                                          m2l_cap, stab, nsrc);
}
#undef EB_XLAT_ARGS

// Several consecutive levels of M2M (upward) or L2L (downward) in one launch, for the coarse
// levels, whose single launches are bound by launch latency (about 10 us each for a few dozen
// to a few thousand boxes): one block per box of level `root`, which runs the translations
// of its subtree level by level with a barrier between levels.  Each wave takes four targets
// of one parity class at a time (the L2L tables differ by class; all four classes' rows are
// in LDS), through the same per-wave product as the one-level kernels (xlat_wave).
struct XlatSub {
  int root;            // level of the subtree roots: the grid is 4^root blocks
  int nlev;            // levels run, in order (M2M: finest first, ending at root; L2L: root+1 on)
  int L;               // finest level of the plan (box_range)
  int lev[4];
  int Wt[4], Ws[4], NBPs[4];
  long long offt[4], offs[4];
};

template <int MODE, int KMAX, int KS, int MT, int TR>
__global__ __launch_bounds__(256) void k_xlat_sub(XlatSub X, const int* __restrict__ leaf_start,
                                                  const double2* src, double2* dst,
                                                  const double* __restrict__ Mat, const double2* __restrict__ pre,
                                                  const double2* __restrict__ post) {
  static_assert(MODE != XL_M2L, "interaction lists run on k_xlat");
  // src and dst are one array (a level reads what the level before it wrote, after the
  // barrier), so neither is declared __restrict__
  constexpr int NM = MT * 16 + TR, NK = KS * 4;
  constexpr int CS = TR ? (NM + 16) / 32 * 32 + 16 : NM + 4;
  __shared__ __align__(16) double Cs[NK * CS];
  __shared__ double2 Ps[4 * NK * 2];  // rows: M2M the child slot o, L2L the target's class
  __shared__ double2 Qs[4 * NM];
  for (int i = threadIdx.x; i < NK * NM; i += blockDim.x) {
    const int k = i / NM, m = i % NM;
    Cs[k * CS + m] = Mat[k * KMAX + m];
  }
  for (int i = threadIdx.x; i < 4 * NK; i += blockDim.x) {
    const int row = i / NK, k = i % NK;
    const double2 w = pre[(size_t)row * KMAX + k];
    Ps[2 * i] = make_double2(w.x, -w.y);
    Ps[2 * i + 1] = make_double2(w.y, w.x);
  }
  for (int i = threadIdx.x; i < 4 * NM; i += blockDim.x) {
    const int row = i / NM, m = i % NM;
    Qs[i] = post[(size_t)row * KMAX + m];
  }
  __syncthreads();
  const int side0 = 1 << X.root;
  const int rx = blockIdx.x % side0, ry = blockIdx.x / side0;
  const int wv = threadIdx.x >> 6, t = (threadIdx.x & 15) >> 2;
  for (int e = 0; e < X.nlev; ++e) {
    const int l = X.lev[e];
    const int sd = 1 << (l - X.root);           // the subtree's side at level l, in boxes
    if constexpr (MODE == XL_M2M && MT > 1) {
      // M2M's tables do not depend on the parity class, so the level's sd^2 <= 4 targets
      // (the four boxes below the root, or the root) are one wave's target set: its MT
      // output tiles go to MT waves, whose chains are then one tile long instead of MT
      if (sd <= 2) {
        if (wv < MT) {
          const bool valid = t < sd * sd;
          const int ix = rx * sd + (valid ? t % sd : 0), iy = ry * sd + (valid ? t / sd : 0);
          const int pt = pidx(ix, iy, X.Wt[e]);
          const int Ws = X.Ws[e];
          auto src_box = [&](int o) -> int { return pidx(2 * ix + (o & 1), 2 * iy + (o >> 1), Ws); };
          xlat_wave<MODE, KMAX, KS, MT, TR, 1>(Cs, Ps, Qs, pt, valid, src_box, src + X.offs[e], X.NBPs[e],
                                               dst + X.offt[e], wv, wv == 0);
        }
        __syncthreads();
        continue;
      }
    }
    const int hs = sd >= 2 ? sd / 2 : 1;
    const int nper = hs * hs;                   // targets per class (the root alone when sd = 1)
    const int jpc = (nper + 3) / 4;             // wave jobs per class
    for (int job = wv; job < 4 * jpc; job += 4) {
      const int c = job / jpc, j = (job % jpc) * 4 + t;
      int ix, iy;
      bool valid;
      if (sd >= 2) {
        valid = j < nper;
        const int jj = valid ? j : 0;
        ix = rx * sd + 2 * (jj % hs) + (c & 1);
        iy = ry * sd + 2 * (jj / hs) + (c >> 1);
      } else {  // M2M into the root: the job of its own class
        valid = j == 0 && c == ((rx & 1) | ((ry & 1) << 1));
        ix = rx;
        iy = ry;
      }
      bool live = valid;
      if (MODE == XL_L2L && valid) {  // as the one-level kernel: empty boxes' locals are never read
        int bs, be;
        box_range(leaf_start, X.L, l, ix, iy, bs, be);
        live = be > bs;
      }
      if (__ballot(live) == 0) continue;  // no barrier inside the level's loop
      const int pt = pidx(ix, iy, X.Wt[e]);
      const int Ws = X.Ws[e];
      auto src_box = [&](int o) -> int {
        if (MODE == XL_M2M) return pidx(2 * ix + (o & 1), 2 * iy + (o >> 1), Ws);
        return pidx(ix >> 1, iy >> 1, Ws);
      };
      const double2* Pw = MODE == XL_M2M ? Ps : Ps + (size_t)c * NK * 2;
      const double2* Qw = MODE == XL_M2M ? Qs : Qs + (size_t)c * NM;
      xlat_wave<MODE, KMAX, KS, MT, TR>(Cs, Pw, Qw, pt, live, src_box, src + X.offs[e], X.NBPs[e], dst + X.offt[e]);
    }
    __syncthreads();  // the level's expansions are the next level's sources
  }
}

// L2P at the finest level (SLLVALDR, cadapdrop2.f:331-348):
//   GW_far = phi + sgn (z conj(phi') + conj(psi))
// With PARENT the leaf's own expansion holds only its interaction-list terms, and the
// parent's complete local expansion is evaluated at the point as well: the same sum as a
// level-L L2L followed by one evaluation (the L2L shift is exact), without writing the
// shifted expansions of every leaf.
//
// Work items: the points under one level-(L-1) box (four sibling leaves and their parent),
// ceil(P / L2P_T) items of nearly equal ranges of sorted points; item = (parent Morton key,
// first point, end point).  One wave per item stages the five local expansions it needs into
// LDS with coalesced loads, then lane i evaluates point i from LDS: the ~20 points of a leaf
// share one set of coefficient loads instead of each lane loading all 4p rows itself.
// The near field (k_p2p) has already written RSCAL * (its sum) to out[perm[i]]; this kernel
// adds RSCAL * (the far field) there, so the near field can run beside the far-field passes.
constexpr int L2P_T = 64;

static size_t l2p_item_cap(int n, int L) { return ((size_t)1 << (2 * (L - 1))) + (size_t)n / L2P_T + 1; }

template <int KMAX, bool PARENT, bool TRANS>
__global__ __launch_bounds__(L2P_T) void k_l2p(int L, int p, const int4* __restrict__ items,
                                               const int* __restrict__ nitems, const uint32_t* __restrict__ keys,
                                               const double2* __restrict__ zs, const double2* __restrict__ lo, int W,
                                               const double2* __restrict__ lop, int Wp, OutForm of,
                                               const int32_t* __restrict__ perm, const Geom* __restrict__ g,
                                               double2* __restrict__ out) {
  // the four leaves' locals (child c = x bit | y bit << 1) and the parent's, both series in a
  // row padded by two entries: the lanes of a wave read four leaves' rows at once, and rows
  // 2 KMAX entries apart (384 dwords at K = 48) would all start on the same LDS bank
  __shared__ double2 C[5][2 * KMAX + 2];
  if ((int)blockIdx.x >= *nitems) return;
  const int4 it = items[blockIdx.x];
  const int px = (int)compact2((uint32_t)it.x), py = (int)compact2((uint32_t)it.x >> 1);
  // lane `row` stages that row of each box's two series: all loads first, then all stores
  static_assert(KMAX <= L2P_T, "one lane per expansion row");
  constexpr int NB = PARENT ? 5 : 4;
  const int row = threadIdx.x;
  double2 v[NB][2];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const double2* src = b < 4 ? lo + (size_t)pidx(2 * px + (b & 1), 2 * py + (b >> 1), W) * 2 * KMAX
                               : lop + (size_t)pidx(px, py, Wp) * 2 * KMAX;
#pragma unroll
    for (int sr = 0; sr < 2; ++sr) v[b][sr] = row < p ? src[sr * KMAX + row] : make_double2(0, 0);
  }
  if (row < p) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int sr = 0; sr < 2; ++sr) C[b][sr * KMAX + row] = v[b][sr];
  }
  __syncthreads();
  const int i = it.y + (int)threadIdx.x;
  if (i >= it.z) return;
  const uint32_t key = keys[i];
  const int ix = (int)compact2(key), iy = (int)compact2(key >> 1);
  const int c = (int)(key & 3u);
  const int nside = 1 << L;
  const double h = 64.0 / nside, ir = 2.0 / h;
  const double2 z = zs[i];
  const double2 u = make_double2((z.x - (ix + 0.5) * h) * ir, (z.y - (iy + 0.5) * h) * ir);
  const double hp = 2.0 * h, irp = 2.0 / hp;
  const double2 up = make_double2((z.x - (px + 0.5) * hp) * irp, (z.y - (py + 0.5) * hp) * irp);
  double2 phi = make_double2(0, 0), dphi = make_double2(0, 0), psi = make_double2(0, 0);
  double2 f = make_double2(0, 0), df = make_double2(0, 0), gq = make_double2(0, 0);
  // both Horner chains in one loop; rows p..KMAX-1 of a local expansion are zero
#pragma unroll 4
  for (int m = p - 1; m >= 0; --m) {
    dphi = cmuladd(dphi, u, phi);
    phi = cmuladd(phi, u, C[c][m]);
    psi = cmuladd(psi, u, C[c][KMAX + m]);
    if (PARENT) {
      df = cmuladd(df, up, f);
      f = cmuladd(f, up, C[4][m]);
      gq = cmuladd(gq, up, C[4][KMAX + m]);
    }
  }
  double2 phip = cscale(dphi, ir);
  if (PARENT) {
    phi = cadd(phi, f);
    psi = cadd(psi, gq);
    phip = cadd(phip, cscale(df, irp));
  }
  // z conj(phi') + conj(psi)
  const double2 zc = cmul(z, cconj(phip));
  const double2 t = make_double2(zc.x + psi.x, zc.y - psi.y);
  const double rs = out_scale(of, g, key, L);
  const int o = perm[i];
  const double2 nearv = out[o];
  const double2 fv = out_form<TRANS>(of, phi, t, TRANS ? of.w[i] : make_double2(0, 0));
  out[o] = make_double2(nearv.x + fv.x * rs, nearv.y + fv.y * rs);
}


// Near-field work items: a leaf with T points becomes ceil(T / tmax) items of nearly equal
// target ranges (all with the leaf's whole 3x3 neighbourhood as sources), so that the lanes of
// a block stay busy on large leaves.  item = (Morton key, first target, end target).  tmax is
// the plan's tile (eb_fmm_set_p2p_tile), P2P_TMAX_DEF by default.
constexpr int P2P_TMAX_DEF = 42, P2P_TMAX_MIN = 16, P2P_TMAX_MAX = 256;

// The near-field and the far-field items in one count and one write launch over one scan:
// entries [0, nleaf) count the near-field items of each leaf, [nleaf, nleaf + npar) the
// far-field (L2P) items of each level-(L-1) box; off is their exclusive scan (nleaf + npar + 1).
// Entry t of both is handled by thread t of the blocks after the M2L list's (k_tree_count,
// k_tree_write).
__device__ __forceinline__ void items_count_entry(int t, int nleaf, int npar, const int* __restrict__ leaf_start,
                                                  int tmax, int* __restrict__ cnt) {
  if (t < nleaf) {
    const int T = leaf_start[t + 1] - leaf_start[t];
    cnt[t] = (T + tmax - 1) / tmax;
  } else if (t < nleaf + npar) {
    const int m = t - nleaf;
    const int P = leaf_start[(m + 1) << 2] - leaf_start[m << 2];
    cnt[t] = (P + L2P_T - 1) / L2P_T;
  }
}

__device__ __forceinline__ void items_write_entry(int t, int nleaf, int npar, const int* __restrict__ leaf_start,
                                                  const int* __restrict__ off, int4* __restrict__ p2p_items,
                                                  int4* __restrict__ l2p_items, int* __restrict__ l2p_total) {
  if (t < nleaf) {
    const int T = leaf_start[t + 1] - leaf_start[t];
    const int o = off[t], np = off[t + 1] - o;
    for (int q = 0; q < np; ++q)
      p2p_items[o + q] = make_int4(t, (int)((long)T * q / np), (int)((long)T * (q + 1) / np), 0);
  } else if (t < nleaf + npar) {
    const int m = t - nleaf;
    const int s = leaf_start[m << 2], P = leaf_start[(m + 1) << 2] - s;
    const int o = off[t] - off[nleaf], np = off[t + 1] - off[t];
    for (int q = 0; q < np; ++q)
      l2p_items[o + q] = make_int4(m, s + (int)((long)P * q / np), s + (int)((long)P * (q + 1) / np), 0);
    if (m == 0) *l2p_total = off[nleaf + npar] - off[nleaf];
  }
}

// The tree's two list builds in one count and one write launch of 1024-thread blocks: blocks
// [0, nm2l) (nm2l = 4 * m2l_nb, or 0 without M2L lists) count or write the M2L target lists,
// the blocks after them the work items.
struct TreeLists {
  int nleaf, npar, tmax;  // work items: leaves, level-(L-1) boxes (0 without L2P), tile
  int L, m2l_nb, lmin, nm2l, m2l_cap;
  const int* leaf_start;
  int *cnt, *off;  // item counts and their exclusive scan
  int4 *p2p_items, *l2p_items;
  int* l2p_total;
  int *ccnt, *m2l_list, *m2l_count;
};

__global__ __launch_bounds__(1024) void k_tree_count(TreeLists a) {
  if ((int)blockIdx.x < a.nm2l) {
    m2l_count_block(blockIdx.x, a.L, a.m2l_nb, a.lmin, a.leaf_start, a.ccnt);
    return;
  }
  const int t = ((int)blockIdx.x - a.nm2l) * 1024 + (int)threadIdx.x;
  items_count_entry(t, a.nleaf, a.npar, a.leaf_start, a.tmax, a.cnt);
}

__global__ __launch_bounds__(1024) void k_tree_write(TreeLists a) {
  if ((int)blockIdx.x < a.nm2l) {
    m2l_write_block(blockIdx.x, a.L, a.m2l_nb, a.lmin, a.leaf_start, a.ccnt, a.m2l_cap, a.m2l_list, a.m2l_count);
    return;
  }
  const int t = ((int)blockIdx.x - a.nm2l) * 1024 + (int)threadIdx.x;
  items_write_entry(t, a.nleaf, a.npar, a.leaf_start, a.off, a.p2p_items, a.l2p_items, a.l2p_total);
}

// Near field at the finest level (DONN0DR, cadapdrop2.f:623-722) plus the final scaling
// GW *= RSCAL (cadapdrop.f:370-373) and the scatter back to the caller's order.
// One 64-lane block per target leaf.  The sources of the 3x3 neighbourhood (MKNBOR,
// cadap2.f:705-742) are staged through LDS in chunks of P2P_CH points; with T < 64 targets
// the lanes split into G = 64/T source slices per target, summed through LDS at the end.
constexpr int P2P_CH = 128;  // 6 KB of LDS per block: LDS does not limit the waves per CU

template <bool TRANS>
__device__ __forceinline__ void p2p_item(int itemid, int L, const int* __restrict__ leaf_start,
                                         const int4* __restrict__ items, const double2* __restrict__ zs,
                                         const double2* __restrict__ qa, const double2* __restrict__ qb,
                                         const double2* __restrict__ gwf, const int32_t* __restrict__ perm,
                                         const Geom* __restrict__ g, const OutForm& of, double2* __restrict__ out,
                                         double2* sz, double2* sa, double2* sq) {
  double2* const red = sz;  // the per-slice partial sums reuse the staging buffers
  double2* const red2 = sa;
  static_assert(P2P_CH >= 64, "reduction buffers alias the staging buffers");
  const int nside = 1 << L;
  const int4 item = items[itemid];  // (leaf Morton key, first and end target within the leaf)
  const uint32_t mk = (uint32_t)item.x;
  const int ix = (int)compact2(mk), iy = (int)compact2(mk >> 1);
  const int ls = leaf_start[mk], TL = leaf_start[mk + 1] - ls;  // the whole leaf: its own sources
  const int ts = ls + item.y, te = ls + item.z;                  // this block's targets
  const int T = te - ts;
  const int lane = threadIdx.x;
  // The 3x3 neighbourhood in MKNBOR order (column-outer, row-inner, cadap2.f:734-740) as nine
  // fixed slots, wave-uniform (scalar registers): ck = the points of slots 0..k-1 (S: all
  // nine), dk = slot k's first sorted index minus its first flat position; a slot outside
  // the grid (or in another segment) holds no points.  Named values, not arrays: an array
  // indexed per lane became a waterfall loop over the lanes' slots.
  int c1, c2, c3, c4, c5, c6, c7, c8, S;
  int d0, d1, d2, d3, d4, d5, d6, d7, d8;
  {
    int run = 0;
    auto slot = [&](int k, int& dk) {
      const int jx = ix + k / 3 - 1, jy = iy + k % 3 - 1;
      bool in = jx >= 0 && jy >= 0 && jx < nside && jy < nside;
      if (of.segk) in = in && ((((jx ^ ix) | (jy ^ iy)) >> (L - of.segk)) == 0);
      int s0 = 0, n0 = 0;
      if (in) {
        const uint32_t m = morton((uint32_t)jx, (uint32_t)jy);
        s0 = leaf_start[m];
        n0 = leaf_start[m + 1] - s0;
      }
      dk = __builtin_amdgcn_readfirstlane(s0 - run);
      run += n0;
      return __builtin_amdgcn_readfirstlane(run);
    };
    c1 = slot(0, d0);
    c2 = slot(1, d1);
    c3 = slot(2, d2);
    c4 = slot(3, d3);
    c5 = slot(4, d4);
    c6 = slot(5, d5);
    c7 = slot(6, d6);
    c8 = slot(7, d7);
    S = slot(8, d8);
  }
  const int self0 = c4;
  const double rs = out_scale(of, g, mk, L);
  // each lane owns two consecutive targets (register blocking halves the LDS reads per
  // pair); with TP = ceil(T/2) < 64 target pairs the lanes split into G = 64/TP slices
  const int TP = (T + 1) >> 1;
  const int TPc = TP < 64 ? TP : 64;
  const int G = TP < 64 ? 64 / TP : 1;
  const int tp = lane % TPc, gs = lane / TPc;
  for (int t0 = 0; t0 < T; t0 += 128) {
    const int i0 = ts + t0 + 2 * tp, i1 = i0 + 1;
    const bool act = gs < G && t0 + 2 * tp < T;
    const bool has1 = act && i1 < te;
    const double2 z0 = act ? zs[i0] : make_double2(0, 0);
    const double2 z1 = has1 ? zs[i1] : z0;
    const int fself0 = self0 + (ts - ls) + t0 + 2 * tp;  // flat index of point i0 among the sources
    const int fself1 = has1 ? fself0 + 1 : -1;
    double2 a10 = make_double2(0, 0), a20 = make_double2(0, 0);
    double2 a11 = make_double2(0, 0), a21 = make_double2(0, 0);
    for (int c0 = 0; c0 < S; c0 += P2P_CH) {
      const int n = min(P2P_CH, S - c0);
      __syncthreads();
      // each lane stages flat positions lane and lane + 64 of the chunk; the slot of a flat
      // position is the last slot starting at or before it (empty slots are passed over)
      static_assert(P2P_CH == 128, "two staged points per lane");
      // a sum of steps rather than a select chain, which the compiler would turn into a
      // table lookup through private memory
      auto src_of = [&](int flat) {
        int j = flat + d0;
        j += flat >= c1 ? d1 - d0 : 0;
        j += flat >= c2 ? d2 - d1 : 0;
        j += flat >= c3 ? d3 - d2 : 0;
        j += flat >= c4 ? d4 - d3 : 0;
        j += flat >= c5 ? d5 - d4 : 0;
        j += flat >= c6 ? d6 - d5 : 0;
        j += flat >= c7 ? d7 - d6 : 0;
        j += flat >= c8 ? d8 - d7 : 0;
        return j;
      };
      const bool u0 = lane < n, u1 = lane + 64 < n;
      const int j0 = u0 ? src_of(c0 + lane) : 0, j1 = u1 ? src_of(c0 + lane + 64) : 0;
      double2 gz0, ga0, gq0, gz1, ga1, gq1;
      if (u0) {
        gz0 = zs[j0];
        ga0 = qa[j0];
        gq0 = qb[j0];
      }
      if (u1) {
        gz1 = zs[j1];
        ga1 = qa[j1];
        gq1 = qb[j1];
      }
      if (u0) {
        sz[lane] = gz0;
        sa[lane] = ga0;
        sq[lane] = gq0;
      }
      if (u1) {
        sz[lane + 64] = gz1;
        sa[lane + 64] = ga1;
        sq[lane + 64] = gq1;
      }
      __syncthreads();
      if (act) {
        // chunk positions [sa0, sb0) hold the leaf's own points: only there can a source be
        // the target itself, so only that range runs the masked pair
        const int fs0 = fself0 - c0, fs1 = fself1 - c0;
        const int sa0 = min(max(self0 - c0, 0), n), sb0 = min(max(self0 + TL - c0, 0), n);
        // this slice's positions gs, gs + G, ... in three runs: before, inside and after the
        // leaf's own points
        int f = gs;
#pragma unroll 2
        for (; f < sa0; f += G) {
          const double2 zj = sz[f], qj = sa[f], bj = sq[f];
          p2p_pair<false, true>(z0, zj, qj, bj, false, a10, a20);
          p2p_pair<false, true>(z1, zj, qj, bj, false, a11, a21);
        }
#pragma unroll 2
        for (; f < sb0; f += G) {
          const double2 zj = sz[f], qj = sa[f], bj = sq[f];
          p2p_pair<true, true>(z0, zj, qj, bj, f == fs0, a10, a20);
          p2p_pair<true, true>(z1, zj, qj, bj, f == fs1, a11, a21);
        }
#pragma unroll 2
        for (; f < n; f += G) {
          const double2 zj = sz[f], qj = sa[f], bj = sq[f];
          p2p_pair<false, true>(z0, zj, qj, bj, false, a10, a20);
          p2p_pair<false, true>(z1, zj, qj, bj, false, a11, a21);
        }
      }
    }
    __syncthreads();
    // the output form is real-linear: applied per slice before the slices' sum
    const double2 zero = make_double2(0, 0);
    red[lane] = out_form<TRANS>(of, a10, a20, TRANS && act ? of.w[i0] : zero);
    red2[lane] = out_form<TRANS>(of, a11, a21, TRANS && has1 ? of.w[i1] : zero);
    __syncthreads();
    if (lane < TPc && t0 + 2 * lane < T) {
      const int j0 = ts + t0 + 2 * lane, j1 = j0 + 1;
      double2 s0 = gwf ? gwf[j0] : make_double2(0, 0);
      for (int q = 0; q < G; ++q) s0 = cadd(s0, red[lane + q * TPc]);
      out[perm[j0]] = make_double2(s0.x * rs, s0.y * rs);
      if (j1 < te) {
        double2 s1 = gwf ? gwf[j1] : make_double2(0, 0);
        for (int q = 0; q < G; ++q) s1 = cadd(s1, red2[lane + q * TPc]);
        out[perm[j1]] = make_double2(s1.x * rs, s1.y * rs);
      }
    }
    __syncthreads();
  }
}

// One block per item (the grid is the item capacity; blocks past the count leave).  A smaller
// persistent grid striding over the items, meant to leave room for the far-field passes running
// beside it, measured slower at every size from 1024 to 8192 blocks.
template <bool TRANS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) void k_p2p(
    int L, const int* __restrict__ leaf_start, const int4* __restrict__ items, const int* __restrict__ nitems,
    const double2* __restrict__ zs, const double2* __restrict__ qa, const double2* __restrict__ qb,
    const double2* __restrict__ gwf, const int32_t* __restrict__ perm, const Geom* __restrict__ g, OutForm of,
    double2* __restrict__ out) {
  __shared__ double2 sz[P2P_CH], sa[P2P_CH], sq[P2P_CH];
  const int ni = *nitems;
  for (int it = blockIdx.x; it < ni; it += gridDim.x)
    p2p_item<TRANS>(it, L, leaf_start, items, zs, qa, qb, gwf, perm, g, of, out, sz, sa, sq);
}

// O(N^2) direct evaluation of the same kernel (small N).  Grid (target tiles of 64, source
// chunks): each one-wave block sums one chunk of sources for 64 targets, staged through LDS 64
// at a time, and writes its partial (qa/d and the bracket, separately); k_direct_sum adds the
// chunks in a fixed order.  A few thousand points thus fill the chip instead of n / 256
// workgroups of the one-pass form.  The reciprocal keeps two Newton steps.
constexpr int DIR_T = 64;
constexpr int DIR_MAXBLK = 4096;  // target tiles x chunks at most (fewer, longer chunks above)

__global__ __launch_bounds__(DIR_T) void k_direct_part(int n, int sch, const double2* __restrict__ zs,
                                                       const double2* __restrict__ qa, const double2* __restrict__ qb,
                                                       const int* __restrict__ guard, double2* __restrict__ part) {
  if (guard && !*guard) return;  // the solver has stopped iterating: the product is not used
  __shared__ double2 sz[DIR_T], sa[DIR_T], sbq[DIR_T];
  const int i = blockIdx.x * DIR_T + threadIdx.x;
  const int c = blockIdx.y;
  const int j0 = c * sch, j1 = min(n, j0 + sch);
  const double2 zi = i < n ? zs[i] : make_double2(0, 0);
  double2 acc1 = make_double2(0, 0), acc2 = make_double2(0, 0);
  for (int t0 = j0; t0 < j1; t0 += DIR_T) {
    const int j = t0 + threadIdx.x;
    if (j < j1) {
      sz[threadIdx.x] = zs[j];
      sa[threadIdx.x] = qa[j];
      sbq[threadIdx.x] = qb[j];
    }
    __syncthreads();
    const int m = min(DIR_T, j1 - t0);
    if (i < n) {
      if (i >= t0 && i < t0 + m) {  // this tile holds the target itself
        for (int jj = 0; jj < m; ++jj) p2p_pair<true, false>(zi, sz[jj], sa[jj], sbq[jj], t0 + jj == i, acc1, acc2);
      } else {
        for (int jj = 0; jj < m; ++jj) p2p_pair<false, false>(zi, sz[jj], sa[jj], sbq[jj], false, acc1, acc2);
      }
    }
    __syncthreads();
  }
  if (i < n) {
    part[2 * ((size_t)c * n + i)] = acc1;
    part[2 * ((size_t)c * n + i) + 1] = acc2;
  }
}

template <bool TRANS>
__global__ __launch_bounds__(256) void k_direct_sum(int n, int nch, const double2* __restrict__ part,
                                                    const int32_t* __restrict__ perm, const Geom* __restrict__ g,
                                                    OutForm of, double2* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 acc1 = make_double2(0, 0), acc2 = make_double2(0, 0);
  for (int c = 0; c < nch; ++c) {
    acc1 = cadd(acc1, part[2 * ((size_t)c * n + i)]);
    acc2 = cadd(acc2, part[2 * ((size_t)c * n + i) + 1]);
  }
  const double rs = g->rscal;
  const double2 v = out_form<TRANS>(of, acc1, acc2, TRANS ? of.w[i] : make_double2(0, 0));
  // perm == nullptr: the points are in the caller's order
  out[perm ? perm[i] : i] = make_double2(v.x * rs, v.y * rs);
}

// chunk length (a multiple of DIR_T) and count for n points: about DIR_MAXBLK blocks in all
static void direct_chunks(int n, int& sch, int& nch) {
  const int tiles = (n + DIR_T - 1) / DIR_T;
  int want = std::max(1, std::min(tiles, DIR_MAXBLK / tiles));
  sch = ((n + want - 1) / want + DIR_T - 1) / DIR_T * DIR_T;
  nch = (n + sch - 1) / sch;
}

// ------------------------------------------------------------------------------------------
// The direct (O(N^2)) evaluation without allocation: eb_direct_apply_noalloc.  The plan path
// (eb_fmm_plan_create in EB_MODE_DIRECT + eb_fmm_apply) allocates its bounding-box partials,
// sorted points, charges and per-chunk partial sums with hipMalloc; this entry point runs the
// same kernels in the caller's point order with those buffers carved from `scratch`, a
// file-scope device array of fixed size that each device holds until the process ends.
// ------------------------------------------------------------------------------------------
// Largest part of the scratch: 2 x nch x n partial sums.  direct_chunks keeps (target tiles of
// DIR_T) x chunks <= DIR_MAXBLK, so nch x n <= DIR_MAXBLK x DIR_T for every n.
constexpr size_t NA_PART = (size_t)2 * DIR_MAXBLK * DIR_T;            // double2 elements
constexpr size_t NA_PTS = (size_t)EB_DIRECT_MAX;                        // points
constexpr size_t NA_BB = 1024;                                          // bounding-box partials
constexpr size_t NA_BYTES = 256 + NA_BB * sizeof(double4) + 3 * NA_PTS * sizeof(double2) + NA_PART * sizeof(double2);
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// scaled coordinates and charges in the caller's order (k_keys' scaling, k_charges' charges)
__global__ void k_na_prep(int n, const double* __restrict__ x, const double* __restrict__ y,
                          const double2* __restrict__ wdz, const double2* __restrict__ dens, const Geom* __restrict__ g,
                          double sqb, int hilbert, double2* __restrict__ zs, double2* __restrict__ qa,
                          double2* __restrict__ qb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double rs = g->rscal;
  zs[i] = make_double2((x[i] - g->rminx) * rs, (y[i] - g->rminy) * rs);
  const double2 d = dens[i];
  const double2 dw = cmul(d, wdz[i]);
  if (hilbert) {
    qa[i] = make_double2(-dw.x, -dw.y);
    qb[i] = make_double2(0.0, 0.0);
    return;
  }
  const double c = -0.15915494309189533577;  // -1/(2 pi): 1/(2 pi i) = -i/(2 pi)
  qa[i] = make_double2(-c * dw.y, c * dw.x);
  const double2 w = wdz[i];
  const double2 t = cadd(cscale(cmul(cconj(d), w), sqb), cmul(d, cconj(w)));
  qb[i] = make_double2(-c * t.y, c * t.x);
}

namespace {
std::mutex na_mu;
constexpr int NA_MAXDEV = 64;
void* na_ptr[NA_MAXDEV] = {};  // device address of `scratch` on each device, once looked up
}  // namespace

// hipGetSymbolAddress once per device, with that device current (the caller's current device)
static int na_scratch(unsigned char** out) {
  int dev = 0;
  EB_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NA_MAXDEV) {
    eb_set_error("eb_direct_apply_noalloc: device %d beyond the %d the library tracks", dev, NA_MAXDEV);
    return EB_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(na_mu);
  if (!na_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      eb_set_error("eb_direct_apply_noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                   hipGetErrorString(e));
      return EB_ERR_HIP;
    }
    na_ptr[dev] = p;
  }
  *out = (unsigned char*)na_ptr[dev];
  return EB_OK;
}

extern "C" int eb_direct_apply_noalloc(int32_t n, const double* x, const double* y, const double* wdz,
                                       const double* dens, int32_t variant, double* gradw, void* stream) {
  if (n < 2 || n > EB_DIRECT_MAX || !x || !y || !wdz || !dens || !gradw) {
    eb_set_error("eb_direct_apply_noalloc: need 2 <= n <= %d points (its fixed scratch) and non-null arrays (n=%d)",
                 EB_DIRECT_MAX, n);
    return EB_ERR_ARG;
  }
  if (variant < 0 || variant > 3) {
    eb_set_error("eb_direct_apply_noalloc: bad variant %d", variant);
    return EB_ERR_ARG;
  }
  unsigned char* base = nullptr;
  if (int rc = na_scratch(&base)) return rc;
  int sch = 0, nch = 0;
  direct_chunks(n, sch, nch);
  const int bbblocks = std::min((int)NA_BB, (n + BB_THREADS - 1) / BB_THREADS);
  if ((size_t)2 * nch * n > NA_PART) {  // cannot happen by direct_chunks' bound; kept as a guard
    eb_set_error("eb_direct_apply_noalloc: %d chunks of %d points exceed the scratch", nch, n);
    return EB_ERR_ARG;
  }
  // scratch layout: Geom | bounding-box partials | zs, qa, qb (n each) | partial sums
  Geom* g = (Geom*)base;
  double4* bb = (double4*)(base + 256);
  double2* zs = (double2*)(base + 256 + NA_BB * sizeof(double4));
  double2* qa = zs + NA_PTS;
  double2* qb = qa + NA_PTS;
  double2* part = qb + NA_PTS;
  const double sgn = (variant == EB_VARIANT_DR) ? -1.0 : (variant == EB_VARIANT_HILBERT ? 0.0 : 1.0);
  const double sqb = (variant == EB_VARIANT_PLUS) ? 1.0 : -1.0;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_bbox_part, dim3(bbblocks), dim3(BB_THREADS), 0, st, x, y, n, bb);
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(BB_THREADS), 0, st, bb, bbblocks, g);
  hipLaunchKernelGGL(k_na_prep, dim3((n + 255) / 256), dim3(256), 0, st, n, x, y, (const double2*)wdz,
                     (const double2*)dens, g, sqb, variant == EB_VARIANT_HILBERT ? 1 : 0, zs, qa, qb);
  hipLaunchKernelGGL(k_direct_part, dim3((n + DIR_T - 1) / DIR_T, nch), dim3(DIR_T), 0, st, n, sch, zs, qa, qb,
                     (const int*)nullptr, part);
  hipLaunchKernelGGL(k_direct_sum<false>, dim3((n + 255) / 256), dim3(256), 0, st, n, nch, part,
                     (const int32_t*)nullptr, g, OutForm{sgn, sqb, OUT_FORWARD, nullptr}, (double2*)gradw);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

// ------------------------------------------------------------------------------------------
// The O(N^2) sum at a subset of targets (eb_direct_subset): every source against each listed
// target, in the caller's coordinates (the kernel is translation invariant and scales as
// 1/length, so no RSCAL is needed), with the charges formed from the density as k_charges
// forms them.  Nothing of the tree, the lists or the expansions enters: it is the checker of
// accuracy measurements (the autotuner, the benchmark), about 1 ms for 1024 targets of 2^20
// points.  Grid (target tiles of DS_T) x (source chunks of DS_CH); the chunk partials are summed
// in a fixed order.
// ------------------------------------------------------------------------------------------
constexpr int DS_T = 64, DS_CH = 4096;

__global__ __launch_bounds__(DS_T) void k_dsub_part(int n, const double* __restrict__ x, const double* __restrict__ y,
                                                    const double2* __restrict__ wdz, const double2* __restrict__ dens,
                                                    double sqb, int hilbert, int ntgt, const int* __restrict__ tgt,
                                                    double2* __restrict__ part) {
  __shared__ double2 sz[DS_T], sa[DS_T], sb[DS_T];
  const int t = blockIdx.x * DS_T + threadIdx.x;
  const int c = blockIdx.y;
  const int j0 = c * DS_CH, j1 = min(n, j0 + DS_CH);
  const int i = t < ntgt ? tgt[t] : -1;
  const double2 zi = i >= 0 ? make_double2(x[i], y[i]) : make_double2(0, 0);
  double2 acc1 = make_double2(0, 0), acc2 = make_double2(0, 0);
  for (int s0 = j0; s0 < j1; s0 += DS_T) {
    const int j = s0 + threadIdx.x;
    if (j < j1) {
      sz[threadIdx.x] = make_double2(x[j], y[j]);
      const double2 d = dens[j], w = wdz[j];
      const double2 dw = cmul(d, w);
      if (hilbert) {
        sa[threadIdx.x] = make_double2(-dw.x, -dw.y);
        sb[threadIdx.x] = make_double2(0.0, 0.0);
      } else {
        const double cc = -0.15915494309189533577;  // -1/(2 pi): 1/(2 pi i) = -i/(2 pi)
        sa[threadIdx.x] = make_double2(-cc * dw.y, cc * dw.x);
        const double2 q = cadd(cscale(cmul(cconj(d), w), sqb), cmul(d, cconj(w)));
        sb[threadIdx.x] = make_double2(-cc * q.y, cc * q.x);
      }
    }
    __syncthreads();
    const int m = min(DS_T, j1 - s0);
    if (i >= 0)
      for (int jj = 0; jj < m; ++jj) p2p_pair<true, false>(zi, sz[jj], sa[jj], sb[jj], s0 + jj == i, acc1, acc2);
    __syncthreads();
  }
  if (t < ntgt) {
    part[2 * ((size_t)c * ntgt + t)] = acc1;
    part[2 * ((size_t)c * ntgt + t) + 1] = acc2;
  }
}

__global__ __launch_bounds__(256) void k_dsub_sum(int ntgt, int nch, const double2* __restrict__ part, double sgn,
                                                  double2* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntgt) return;
  double2 a1 = make_double2(0, 0), a2 = make_double2(0, 0);
  for (int c = 0; c < nch; ++c) {
    a1 = cadd(a1, part[2 * ((size_t)c * ntgt + t)]);
    a2 = cadd(a2, part[2 * ((size_t)c * ntgt + t) + 1]);
  }
  out[t] = make_double2(a1.x + sgn * a2.x, a1.y + sgn * a2.y);
}

extern "C" int eb_direct_subset(int32_t n, const double* x, const double* y, const double* wdz, const double* dens,
                                int32_t variant, int32_t ntgt, const int32_t* tgt, double* gradw, void* stream) {
  if (n < 2 || ntgt < 1 || !x || !y || !wdz || !dens || !tgt || !gradw || variant < 0 || variant > 3) {
    eb_set_error("eb_direct_subset: bad arguments (n=%d ntgt=%d variant=%d)", n, ntgt, variant);
    return EB_ERR_ARG;
  }
  const int nch = (n + DS_CH - 1) / DS_CH;
  double2* part = nullptr;
  if (hipMalloc((void**)&part, sizeof(double2) * 2 * (size_t)nch * ntgt) != hipSuccess) {
    eb_set_error("eb_direct_subset: hipMalloc of %d chunk partials failed", nch);
    return EB_ERR_NOMEM;
  }
  hipStream_t st = (hipStream_t)stream;
  const double sgn = (variant == EB_VARIANT_DR) ? -1.0 : (variant == EB_VARIANT_HILBERT ? 0.0 : 1.0);
  const double sqb = (variant == EB_VARIANT_PLUS) ? 1.0 : -1.0;
  hipLaunchKernelGGL(k_dsub_part, dim3((ntgt + DS_T - 1) / DS_T, nch), dim3(DS_T), 0, st, n, x, y,
                     (const double2*)wdz, (const double2*)dens, sqb, variant == EB_VARIANT_HILBERT ? 1 : 0, ntgt, tgt,
                     part);
  hipLaunchKernelGGL(k_dsub_sum, dim3((ntgt + 255) / 256), dim3(256), 0, st, ntgt, nch, part, sgn, (double2*)gradw);
  hipError_t e = hipGetLastError();
  // the partials are freed once the kernels are done
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(part);
  if (e != hipSuccess) {
    eb_set_error("eb_direct_subset: %s", hipGetErrorString(e));
    return EB_ERR_HIP;
  }
  return EB_OK;
}

// ------------------------------------------------------------------------------------------
// host-side tables (long double, then rounded)
// ------------------------------------------------------------------------------------------
typedef long double LD;
struct LC {
  LD x, y;
};
static inline LC lmul(LC a, LC b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }

struct Tables {
  std::vector<double> Mat;    // [3][KMAX][KMAX]  M2L C(m+k,k); M2M C(n,k); L2L C(m,j)  (k-major)
  std::vector<double2> U;     // [4][27][KMAX]  M2L pre  u^(k+1)
  std::vector<double2> NU;    // [4][27][KMAX]  M2L post (-u)^m
  std::vector<double2> PM;    // [4][KMAX]      M2M pre  2^-(k+1) d_c^-k
  std::vector<double2> QM;    // [4][KMAX]      M2M post d_c^n
  std::vector<double2> PL;    // [4][KMAX]      L2L pre  d_c^m
  std::vector<double2> QL;    // [4][KMAX]      L2L post (2 d_c)^-j
  std::vector<int> offs;      // [4][27][2]
};

static LD binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  LD r = 1;
  for (int i = 1; i <= k; ++i) r = r * (LD)(n - k + i) / (LD)i;
  return r;
}

static LC lpow(LC z, int n) {
  LC r = {1, 0};
  for (int i = 0; i < n; ++i) r = lmul(r, z);
  return r;
}

static LC linv(LC z) {
  const LD d = z.x * z.x + z.y * z.y;
  return {z.x / d, -z.y / d};
}

static void build_tables(int p, int KMAX, Tables& T) {
  const size_t KK = (size_t)KMAX * KMAX;
  T.Mat.assign(3 * KK, 0.0);
  for (int k = 0; k < p; ++k)
    for (int m = 0; m < p; ++m) {
      T.Mat[0 * KK + (size_t)k * KMAX + m] = (double)binom(m + k, k);
      T.Mat[1 * KK + (size_t)k * KMAX + m] = (double)binom(m, k);  // [k][n] = C(n,k)
      T.Mat[2 * KK + (size_t)k * KMAX + m] = (double)binom(k, m);  // [m][j] = C(m,j)
    }
  T.U.assign((size_t)4 * 27 * KMAX, make_double2(0, 0));
  T.NU.assign((size_t)4 * 27 * KMAX, make_double2(0, 0));
  T.offs.assign(4 * 27 * 2, 0);
  for (int c = 0; c < 4; ++c) {
    int dxs[27], dys[27];
    int no = m2l_offsets(c & 1, c >> 1, dxs, dys);
    for (int o = 0; o < no; ++o) {
      T.offs[(c * 27 + o) * 2 + 0] = dxs[o];
      T.offs[(c * 27 + o) * 2 + 1] = dys[o];
      // z0 = c_t - c_s = -(dx + i dy) * 2r  ->  u = r / z0 = -1 / (2 (dx + i dy))
      LD den = 2.0L * ((LD)dxs[o] * dxs[o] + (LD)dys[o] * dys[o]);
      LC u = {-(LD)dxs[o] / den, (LD)dys[o] / den};
      LC nu = {-u.x, -u.y};
      LC pw = u, pn = {1, 0};
      for (int k = 0; k < KMAX; ++k) {
        T.U[(size_t)(c * 27 + o) * KMAX + k] = make_double2((double)pw.x, (double)pw.y);
        T.NU[(size_t)(c * 27 + o) * KMAX + k] = make_double2((double)pn.x, (double)pn.y);
        pw = lmul(pw, u);
        pn = lmul(pn, nu);
      }
    }
  }
  T.PM.assign((size_t)4 * KMAX, make_double2(0, 0));
  T.QM = T.PM;
  T.PL = T.PM;
  T.QL = T.PM;
  for (int c = 0; c < 4; ++c) {
    // child centre - parent centre = (+-r_c, +-r_c);  d = that / r_parent = (+-1 +- i)/2
    const LC d = {(c & 1) ? 0.5L : -0.5L, (c >> 1) ? 0.5L : -0.5L};
    const LC di = linv(d);
    const LC d2i = linv(LC{2 * d.x, 2 * d.y});
    for (int k = 0; k < KMAX; ++k) {
      LC a = lpow(di, k);
      const LD h = powl(0.5L, (LD)(k + 1));
      T.PM[(size_t)c * KMAX + k] = make_double2((double)(h * a.x), (double)(h * a.y));
      LC b = lpow(d, k);
      T.QM[(size_t)c * KMAX + k] = make_double2((double)b.x, (double)b.y);
      T.PL[(size_t)c * KMAX + k] = make_double2((double)b.x, (double)b.y);
      LC q = lpow(d2i, k);
      T.QL[(size_t)c * KMAX + k] = make_double2((double)q.x, (double)q.y);
    }
  }
}

// ------------------------------------------------------------------------------------------
// plan
// ------------------------------------------------------------------------------------------
struct eb_fmm_plan_s {
  int n = 0, L = 1, p = 0, kmax = 0, variant = 0, mode = EB_MODE_FMM;
  double sgn = -1.0, sqb = -1.0;
  hipStream_t st = nullptr;
  // device
  Geom* geom = nullptr;
  double4* bbpart = nullptr;
  double2* zs0 = nullptr;
  double2* zs = nullptr;
  double2* wdzs = nullptr;
  double2* qa = nullptr;
  double2* qb = nullptr;
  double2* gwf = nullptr;
  uint32_t* keys = nullptr;
  uint32_t* keys2 = nullptr;
  int32_t* perm = nullptr;
  int32_t* vals2 = nullptr;
  int* sorttmp = nullptr;
  int* leaf_start = nullptr;
  int4* p2p_items = nullptr;  // near-field work items and their count (after the scan total)
  int p2p_tmax = P2P_TMAX_DEF;
  int* p2p_cnt = nullptr;
  int* p2p_off = nullptr;
  int* p2p_tmp = nullptr;
  int4* l2p_items = nullptr;  // far-field evaluation items (FMM mode, L >= 2); their count
  int* l2p_total = nullptr;   // (p2p_cnt / p2p_off hold both item kinds' counts and offsets)
  // direct mode: the rebuild leaves the points in the caller's order (the O(N^2) sum needs no
  // tree); the sorted tree is made when eb_fmm_plan_ntbox or eb_fmm_plan_info asks for it
  bool tree_sorted = true;
  bool hist0_ready = false;  // the radix sort's first histogram of the current keys is in sorttmp
  double2* wdz0 = nullptr;  // direct mode: the weights in the caller's order
  int* ntbox = nullptr;
  double2* dir_part = nullptr;  // direct mode: 2 x n partial sums per source chunk
  int dir_sch = 0, dir_nch = 0;
  const int* guard = nullptr;  // direct mode: device flag; 0 = skip the pair sums (eb_fmm_set_guard)
  int* m2l_list = nullptr;   // compacted M2L targets (k_tree_count, k_tree_write), 4 x m2l_cap
  int* m2l_count = nullptr;  // 4
  int* m2l_ccnt = nullptr;   // per (class, chunk of 1024 boxes) counts, 4 x m2l_nb
  int m2l_cap = 0, m2l_nb = 0;
  double2* mp = nullptr;
  double2* lo = nullptr;
  size_t lvl_off[EB_MAXLEV + 2] = {0};
  int lvl_W[EB_MAXLEV + 2] = {0};
  int lvl_NBP[EB_MAXLEV + 2] = {0};
  double* dMat = nullptr;
  double2 *dU = nullptr, *dNU = nullptr, *dPM = nullptr, *dQM = nullptr, *dPL = nullptr, *dQL = nullptr;
  int* doffs = nullptr;
  // timing
  bool timing = false;
  hipEvent_t ev[8] = {};
  double stage_ms[7] = {0};
  hipEvent_t ev_switch = nullptr;  // orders a new stream after the plan's previous one
  // the near field runs on st2 beside the far-field passes on st (fork after the charges,
  // join before L2P, which adds the far field to the near field's output)
  hipStream_t st2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // segmented plans (eb_fmm_plan_create_segmented): nseg independent point sets, segment s in
  // box s (Morton order) of level segk, each with its own bounding box and scale
  int nseg = 0, segk = 0;
  int* seg_off = nullptr;       // nseg + 1 point offsets
  double4* seg_geo = nullptr;   // per segment: (XMIN, YMIN, scale, 0)
  double* seg_scale = nullptr;  // per segment: the scale (the output factor, as RSCAL)
};

static void plan_free(eb_fmm_plan P) {
  if (!P) return;
  void* ptrs[] = {P->geom, P->bbpart, P->zs0, P->zs, P->wdzs, P->qa, P->qb, P->gwf, P->keys, P->keys2,
                  P->perm, P->vals2, P->sorttmp, P->leaf_start, P->p2p_items, P->p2p_cnt, P->p2p_off,
                  P->p2p_tmp, P->l2p_items, P->l2p_total, P->wdz0, P->ntbox, P->m2l_list, P->m2l_count, P->m2l_ccnt, P->mp, P->lo, P->dMat,
                  P->dU, P->dNU, P->dPM, P->dQM, P->dPL, P->dQL, P->doffs, P->dir_part, P->seg_off, P->seg_geo,
                  P->seg_scale};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  for (auto& e : P->ev)
    if (e) (void)hipEventDestroy(e);
  if (P->ev_switch) (void)hipEventDestroy(P->ev_switch);
  if (P->ev_fork) (void)hipEventDestroy(P->ev_fork);
  if (P->ev_join) (void)hipEventDestroy(P->ev_join);
  if (P->st2) {
    (void)hipStreamSynchronize(P->st2);
    (void)hipStreamDestroy(P->st2);
  }
  delete P;
}

template <typename T>
static int dmalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
  if (e != hipSuccess) {
    eb_set_error("hipMalloc(%zu bytes) failed: %s", sizeof(T) * count, hipGetErrorString(e));
    return EB_ERR_NOMEM;
  }
  return EB_OK;
}

static int pick_kmax(int p) {
  if (p <= 16) return 16;
  if (p <= 32) return 32;
  if (p <= 48) return 48;
  return 64;
}

// Input terms of the M2L's second source group (xlat_wave: KSF = ceil(2 KS / 3) k steps of
// KS = min(K / 4, ceil(p / 4))), at most p
static int m2l_far_terms(int p, int kmax) {
  const int ks = std::min(kmax / 4, (p + 3) / 4);
  return std::min(p, 4 * ((2 * ks + 2) / 3));
}

// Move the plan to `stream` (NULL keeps the current one).  Work already queued on the old stream
// (a tree build, the last apply) is ordered before anything the new stream runs next.
static int use_stream(eb_fmm_plan P, void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (!s || s == P->st) return EB_OK;
  if (!P->ev_switch) EB_HIP_CHECK(hipEventCreateWithFlags(&P->ev_switch, hipEventDisableTiming));
  EB_HIP_CHECK(hipEventRecord(P->ev_switch, P->st));
  EB_HIP_CHECK(hipStreamWaitEvent(s, P->ev_switch, 0));
  P->st = s;
  return EB_OK;
}

#define EB_TRY(x)            \
  do {                       \
    int _rc = (x);           \
    if (_rc != EB_OK) {      \
      plan_free(P);          \
      return _rc;            \
    }                        \
  } while (0)

// Bin the points (tree build): bounding box and RSCAL, scaled coordinates and finest-level
// keys, stable sort by key, leaf table, gather into sorted order, non-empty counts per level.
// most near-field work items for tile tmax: one per non-empty leaf plus one per extra tmax
// points (sum over leaves of ceil(T/tmax) <= leaves + sum (T-1)/tmax <= leaves + n/tmax)
static size_t p2p_item_cap(int n, int L, int tmax) { return ((size_t)1 << (2 * L)) + (size_t)n / tmax + 1; }

// The near-field work items of the current leaves and tile and, in FMM mode, the far-field
// (L2P) items: one count launch, one scan over both, one write launch.  Their counts stay on
// the device (p2p_off[nleaf] and *l2p_total).  With with_m2l the same two launches also build
// the M2L target lists.
static int build_items(eb_fmm_plan P, bool with_m2l) {
  hipStream_t st = P->st;
  TreeLists a{};
  a.nleaf = 1 << (2 * P->L);
  a.npar = P->l2p_items ? 1 << (2 * (P->L - 1)) : 0;
  a.tmax = P->p2p_tmax;
  a.L = P->L;
  a.leaf_start = P->leaf_start;
  a.cnt = P->p2p_cnt;
  a.off = P->p2p_off;
  a.p2p_items = P->p2p_items;
  a.l2p_items = P->l2p_items;
  a.l2p_total = P->l2p_total;
  if (with_m2l && P->m2l_list) {
    // a segmented plan's boxes interact from two levels below the segments' level: the
    // interaction lists of the segment boxes' children lie wholly in other segments
    a.lmin = P->segk ? P->segk + 2 : 2;
    a.m2l_nb = P->m2l_nb;
    a.nm2l = 4 * P->m2l_nb;
    a.m2l_cap = P->m2l_cap;
    a.ccnt = P->m2l_ccnt;
    a.m2l_list = P->m2l_list;
    a.m2l_count = P->m2l_count;
  }
  const int m = a.nleaf + a.npar;
  const int blocks = a.nm2l + (m + 1023) / 1024;
  hipLaunchKernelGGL(k_tree_count, dim3(blocks), dim3(1024), 0, st, a);
  const int rc = scan_exclusive(P->p2p_cnt, P->p2p_off, m, true, P->p2p_tmp, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_tree_write, dim3(blocks), dim3(1024), 0, st, a);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

// The sort of the points by leaf and everything that follows from it (leaf ranges, gathered
// points, work items, the M2L target lists), from the keys and identity permutation of k_keys
// and the points and weights in the caller's order (zs0, wsrc).
static int sort_tree(eb_fmm_plan P, const double2* wsrc) {
  hipStream_t st = P->st;
  const int n = P->n, L = P->L;
  const int nleaf = 1 << (2 * L);
  int rc = radix_sort_pairs(P->keys, P->perm, P->keys2, P->vals2, n, 2 * L, P->sorttmp, st, P->hist0_ready);
  P->hist0_ready = false;
  if (rc) return rc;
  const int t = std::max(n, nleaf + 1);
  hipLaunchKernelGGL(k_leaf_gather, dim3((t + 255) / 256), dim3(256), 0, st, P->keys, n, nleaf, P->leaf_start,
                     P->perm, P->zs0, wsrc, P->zs, P->wdzs);
  if ((rc = build_items(P, true)) != EB_OK) return rc;
  P->tree_sorted = true;
  return EB_OK;
}

// points and weights in the caller's order, identity permutation (direct mode)
__global__ void k_direct_prep(int n, const double2* __restrict__ zs0, const double2* __restrict__ wdz,
                              double2* __restrict__ zs, double2* __restrict__ wdz0, double2* __restrict__ wdzs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  zs[i] = zs0[i];
  const double2 w = wdz[i];
  wdz0[i] = w;
  wdzs[i] = w;
}

// direct mode: the sorted tree on demand (the tables and counts of eb_fmm_plan_ntbox / _info);
// the permutation then applies to the points as before
static int ensure_tree(eb_fmm_plan P) {
  if (P->tree_sorted) return EB_OK;
  return sort_tree(P, P->wdz0);
}

static int build_tree(eb_fmm_plan P, const double* x, const double* y, const double* wdz) {
  hipStream_t st = P->st;
  const int n = P->n, L = P->L;
  const int nleaf = 1 << (2 * L);
  // at most 1024 bounding-box partials: every block of k_keys reduces them all, one per thread
  const int bbblocks = std::min(1024, (n + BB_THREADS - 1) / BB_THREADS);
  if (P->nseg) {
    hipLaunchKernelGGL(k_seg_bbox, dim3(P->nseg), dim3(BB_THREADS), 0, st, P->nseg, P->seg_off, x, y,
                       64.0 / (1 << P->segk), P->seg_geo, P->seg_scale, P->geom);
    hipLaunchKernelGGL(k_seg_keys, dim3((n + 255) / 256), dim3(256), 0, st, n, P->nseg, P->seg_off, x, y, L, P->segk,
                       P->seg_geo, P->zs0, P->keys, P->perm);
  } else {
    hipLaunchKernelGGL(k_bbox_part, dim3(bbblocks), dim3(BB_THREADS), 0, st, x, y, n, P->bbpart);
    const int nblk = (n + RS_TILE_KEYS - 1) / RS_TILE_KEYS;
    hipLaunchKernelGGL(k_keys, dim3(nblk), dim3(1024), 0, st, x, y, n, L, P->bbpart, bbblocks, P->geom, P->zs0,
                       P->keys, P->perm, P->sorttmp, nblk);
  }
  P->hist0_ready = P->nseg == 0;
  EB_HIP_CHECK(hipGetLastError());
  if (P->mode == EB_MODE_DIRECT) {
    // the O(N^2) sum runs on the points in the caller's order: four launches per rebuild
    // instead of the sort, the leaf ranges and the lists
    hipLaunchKernelGGL(k_direct_prep, dim3((n + 255) / 256), dim3(256), 0, st, n, P->zs0, (const double2*)wdz, P->zs,
                       P->wdz0, P->wdzs);
    EB_HIP_CHECK(hipGetLastError());
    P->tree_sorted = false;
    return EB_OK;
  }
  (void)nleaf;
  return sort_tree(P, (const double2*)wdz);
}

// eb_fmm_plan_create (nseg = 0) and eb_fmm_plan_create_segmented (nseg >= 1 segments with the
// host offsets seg_off_h; levmax is then the depth below the segments' level)
static int plan_create(eb_fmm_plan* out, int32_t n, const double* x, const double* y, const double* wdz,
                       int32_t levmax, int32_t nterms, int32_t variant, int32_t mode, void* stream, int32_t nseg,
                       const int32_t* seg_off_h) {
  if (!out) {
    eb_set_error("eb_fmm_plan_create: null plan pointer");
    return EB_ERR_ARG;
  }
  *out = nullptr;
  if (n < 2 || !x || !y || !wdz) {
    eb_set_error("eb_fmm_plan_create: need n >= 2 points and non-null x, y, wdz (n=%d)", n);
    return EB_ERR_ARG;
  }
  int segk = 0;
  if (nseg > 0) {
    while (((int64_t)1 << (2 * segk)) < nseg) ++segk;
    if (mode != EB_MODE_FMM || !seg_off_h || seg_off_h[0] != 0 || seg_off_h[nseg] != n) {
      eb_set_error("eb_fmm_plan_create_segmented: FMM mode and offsets 0 = seg_off[0] <= ... <= seg_off[nseg] = n");
      return EB_ERR_ARG;
    }
    for (int i = 0; i < nseg; ++i)
      if (seg_off_h[i + 1] < seg_off_h[i]) {
        eb_set_error("eb_fmm_plan_create_segmented: seg_off decreases at segment %d", i);
        return EB_ERR_ARG;
      }
    if (levmax < 1 || segk + levmax > EB_MAXLEV) {
      eb_set_error("eb_fmm_plan_create_segmented: %d segments need %d levels above the per-segment levmax=%d; "
                   "the total must stay within 1..%d", nseg, segk, levmax, EB_MAXLEV);
      return EB_ERR_ARG;
    }
    levmax += segk;
  }
  if (levmax < 1 || levmax > EB_MAXLEV) {
    eb_set_error("eb_fmm_plan_create: levmax=%d outside 1..%d", levmax, EB_MAXLEV);
    return EB_ERR_ARG;
  }
  if (nterms < 1 || nterms > 60) {
    eb_set_error("eb_fmm_plan_create: nterms=%d outside 1..60", nterms);
    return EB_ERR_TERMS;
  }
  if (variant < 0 || variant > 3 || (mode != EB_MODE_FMM && mode != EB_MODE_DIRECT)) {
    eb_set_error("eb_fmm_plan_create: bad variant %d or mode %d", variant, mode);
    return EB_ERR_ARG;
  }
  eb_fmm_plan P = new eb_fmm_plan_s();
  P->n = n;
  P->L = levmax;
  P->p = nterms;
  P->kmax = pick_kmax(nterms);
  P->variant = variant;
  P->mode = mode;
  // GW = phi + sgn (z conj(phi') + conj(psi)); the Hilbert product is phi alone
  P->sgn = (variant == EB_VARIANT_DR) ? -1.0 : (variant == EB_VARIANT_HILBERT ? 0.0 : 1.0);
  P->sqb = (variant == EB_VARIANT_PLUS) ? 1.0 : -1.0;
  P->st = (hipStream_t)stream;
  hipStream_t st = P->st;
  const int L = P->L, K = P->kmax;
  const int nleaf = 1 << (2 * L);
  const int bbblocks = std::min(1024, (n + BB_THREADS - 1) / BB_THREADS);
  if (nseg > 0) {
    P->nseg = nseg;
    P->segk = segk;
    EB_TRY(dmalloc(&P->seg_off, nseg + 1));
    EB_TRY(dmalloc(&P->seg_geo, nseg));
    EB_TRY(dmalloc(&P->seg_scale, nseg));
    EB_TRY(hipMemcpyAsync(P->seg_off, seg_off_h, sizeof(int) * (nseg + 1), hipMemcpyHostToDevice, st) == hipSuccess &&
                   hipStreamSynchronize(st) == hipSuccess
               ? EB_OK
               : (eb_set_error("eb_fmm_plan_create_segmented: offset upload failed"), EB_ERR_HIP));
  }

  EB_TRY(dmalloc(&P->geom, 1));
  EB_TRY(dmalloc(&P->bbpart, bbblocks));
  EB_TRY(dmalloc(&P->zs0, n));
  EB_TRY(dmalloc(&P->zs, n));
  EB_TRY(dmalloc(&P->wdzs, n));
  EB_TRY(dmalloc(&P->qa, n));
  EB_TRY(dmalloc(&P->qb, n));
  EB_TRY(dmalloc(&P->keys, n));
  EB_TRY(dmalloc(&P->keys2, n));
  EB_TRY(dmalloc(&P->perm, n));
  EB_TRY(dmalloc(&P->vals2, n));
  EB_TRY(dmalloc(&P->sorttmp, radix_sort_temp_ints(n)));
  EB_TRY(dmalloc(&P->leaf_start, nleaf + 1));
  EB_TRY(dmalloc(&P->p2p_items, p2p_item_cap(n, L, P2P_TMAX_MIN)));
  {
    // near-field counts and, in FMM mode, the far-field items' counts after them
    const int npar = (mode == EB_MODE_FMM && L >= 2) ? 1 << (2 * (L - 1)) : 0;
    EB_TRY(dmalloc(&P->p2p_cnt, nleaf + npar));
    EB_TRY(dmalloc(&P->p2p_off, nleaf + npar + 1));
    EB_TRY(dmalloc(&P->p2p_tmp, scan_temp_ints(nleaf + npar)));
  }
  EB_TRY(dmalloc(&P->ntbox, EB_MAXLEV + 2));
  if (mode == EB_MODE_DIRECT) {
    direct_chunks(n, P->dir_sch, P->dir_nch);
    EB_TRY(dmalloc(&P->dir_part, (size_t)2 * P->dir_nch * n));
    EB_TRY(dmalloc(&P->wdz0, n));
  }
  if (mode == EB_MODE_FMM && L >= 2) {
    const int npar = 1 << (2 * (L - 1));
    EB_TRY(dmalloc(&P->l2p_items, l2p_item_cap(n, L)));
    EB_TRY(dmalloc(&P->l2p_total, 1));
    size_t cap = 0;
    for (int l = 2; l <= L; ++l) cap += ((size_t)1 << (2 * (l - 1))) + 3;
    P->m2l_cap = (int)cap;
    int nb = 0;
    for (int l = 2; l <= L; ++l) nb += ((1 << (2 * (l - 1))) + 1023) >> 10;
    P->m2l_nb = nb;
    EB_TRY(dmalloc(&P->m2l_list, 4 * cap));
    EB_TRY(dmalloc(&P->m2l_count, 4));
    EB_TRY(dmalloc(&P->m2l_ccnt, 4 * (size_t)nb));
  }

  EB_TRY(build_tree(P, x, y, wdz));

  if (mode == EB_MODE_FMM && L >= 2) {
    EB_TRY(hipStreamCreateWithFlags(&P->st2, hipStreamNonBlocking) == hipSuccess &&
                   hipEventCreateWithFlags(&P->ev_fork, hipEventDisableTiming) == hipSuccess &&
                   hipEventCreateWithFlags(&P->ev_join, hipEventDisableTiming) == hipSuccess
               ? EB_OK
               : (eb_set_error("eb_fmm_plan_create: stream or event creation failed"), EB_ERR_HIP));
    size_t off = 0;
    for (int l = 2; l <= L; ++l) {
      const int W = (1 << l) + 6;
      P->lvl_W[l] = W;
      P->lvl_NBP[l] = W * W;
      P->lvl_off[l] = off;
      off += (size_t)2 * K * W * W;
    }
    EB_TRY(dmalloc(&P->mp, off));
    EB_TRY(dmalloc(&P->lo, off));
    // the 3-box halo of every level must read as zero multipoles
    EB_TRY((hipMemsetAsync(P->mp, 0, sizeof(double2) * off, st) == hipSuccess)
               ? EB_OK
               : (eb_set_error("hipMemsetAsync failed"), EB_ERR_HIP));
    // a segmented plan never writes the locals of levels up to segk + 1 (no interaction lists
    // there), which the downward pass reads as zero
    if (P->segk)
      EB_TRY((hipMemsetAsync(P->lo, 0, sizeof(double2) * off, st) == hipSuccess)
                 ? EB_OK
                 : (eb_set_error("hipMemsetAsync failed"), EB_ERR_HIP));
    Tables T;
    build_tables(P->p, K, T);
    EB_TRY(dmalloc(&P->dMat, T.Mat.size()));
    EB_TRY(dmalloc(&P->dU, T.U.size()));
    EB_TRY(dmalloc(&P->dNU, T.NU.size()));
    EB_TRY(dmalloc(&P->dPM, T.PM.size()));
    EB_TRY(dmalloc(&P->dQM, T.QM.size()));
    EB_TRY(dmalloc(&P->dPL, T.PL.size()));
    EB_TRY(dmalloc(&P->dQL, T.QL.size()));
    EB_TRY(dmalloc(&P->doffs, T.offs.size()));
    bool ok = hipMemcpyAsync(P->dMat, T.Mat.data(), sizeof(double) * T.Mat.size(), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(P->dU, T.U.data(), sizeof(double2) * T.U.size(), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(P->dNU, T.NU.data(), sizeof(double2) * T.NU.size(), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(P->dPM, T.PM.data(), sizeof(double2) * T.PM.size(), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(P->dQM, T.QM.data(), sizeof(double2) * T.QM.size(), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(P->dPL, T.PL.data(), sizeof(double2) * T.PL.size(), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(P->dQL, T.QL.data(), sizeof(double2) * T.QL.size(), hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(P->doffs, T.offs.data(), sizeof(int) * T.offs.size(), hipMemcpyHostToDevice, st) == hipSuccess;
    // the host vectors die at scope end: wait for the copies
    ok = ok && hipStreamSynchronize(st) == hipSuccess;
    if (!ok) {
      eb_set_error("eb_fmm_plan_create: table upload failed");
      plan_free(P);
      return EB_ERR_HIP;
    }
  }
  *out = P;
  return EB_OK;
}

extern "C" int eb_fmm_plan_create(eb_fmm_plan* out, int32_t n, const double* x, const double* y,
                                  const double* wdz, int32_t levmax, int32_t nterms, int32_t variant,
                                  int32_t mode, void* stream) {
  return plan_create(out, n, x, y, wdz, levmax, nterms, variant, mode, stream, 0, nullptr);
}

extern "C" int eb_fmm_plan_create_segmented(eb_fmm_plan* out, int32_t n, const double* x, const double* y,
                                            const double* wdz, int32_t nseg, const int32_t* seg_off, int32_t levmax,
                                            int32_t nterms, int32_t variant, void* stream) {
  if (nseg < 1) {
    eb_set_error("eb_fmm_plan_create_segmented: nseg=%d", nseg);
    return EB_ERR_ARG;
  }
  return plan_create(out, n, x, y, wdz, levmax, nterms, variant, EB_MODE_FMM, stream, nseg, seg_off);
}

// Fill the level table of one translation launch: target levels [l0, l1] (one entry each),
// sources at level l + dl.
static XlatLevels xlat_levels(eb_fmm_plan P, int l0, int l1, int dl) {
  XlatLevels LS{};
  LS.L = P->L;
  int tot = 0, e = 0;
  for (int l = l0; l <= l1; ++l, ++e) {
    const int half = 1 << (l - 1);
    LS.lev[e] = l;
    LS.blk0[e] = tot;
    LS.Wt[e] = P->lvl_W[l];
    LS.NBPt[e] = P->lvl_NBP[l];
    LS.offt[e] = (long long)P->lvl_off[l];
    LS.Ws[e] = P->lvl_W[l + dl];
    LS.NBPs[e] = P->lvl_NBP[l + dl];
    LS.offs[e] = (long long)P->lvl_off[l + dl];
    tot += 4 * ((half * half + 15) / 16);  // 16 target boxes per block, per parity class
  }
  LS.nent = e;
  LS.blk0[e] = tot;
  LS.total = tot;
  LS.segk = P->segk;
  return LS;
}

template <int MODE, int K, int KS>
static void launch_xlat_ks(eb_fmm_plan P, const XlatLevels& LS, const double2* src, double2* dst,
                           const double2* pre, const double2* post) {
  // p = 33..36 (K = 48): two MFMA tiles and a 4-row tail on the vector ALUs.  An 8-row tail
  // for p = 37..40 reads its Pascal entries from LDS per source (registers do not hold
  // them), and measured slower than three tiles: 0.84 against 0.68 ms at p = 37.
  constexpr int TR = (K == 48 && KS == 9) ? 4 : 0;
  constexpr int MT = TR ? 2 : K / 16;
  const int grid = (LS.total + 7) / 8 * 8;
  const double* mat = P->dMat + (size_t)MODE * K * K;
  const bool c = MODE == XL_M2L && P->m2l_list != nullptr;
  if constexpr (MODE == XL_M2L) {
    if (P->segk) {
      // a segmented plan: sources outside the target's segment are skipped.  The test takes
      // registers that the two-wave bound of the vector-tail kernel does not leave (it spilled
      // 344 B per lane there), so the tail form runs unbounded too.
      hipLaunchKernelGGL((k_xlat<MODE, K, KS, MT, TR, true>), dim3(grid), dim3(256), 0, P->st, LS, P->leaf_start, src,
                         dst, mat, pre, post, P->doffs, c ? P->m2l_list : nullptr, c ? P->m2l_count : nullptr,
                         P->m2l_cap);
      return;
    }
  }
  if constexpr (MODE == XL_M2L && TR == 4)
    hipLaunchKernelGGL((k_xlat2<MODE, K, KS, MT, TR>), dim3(grid), dim3(256), 0, P->st, LS, P->leaf_start, src, dst,
                       mat, pre, post, P->doffs, c ? P->m2l_list : nullptr, c ? P->m2l_count : nullptr, P->m2l_cap);
  else
    hipLaunchKernelGGL((k_xlat<MODE, K, KS, MT, TR>), dim3(grid), dim3(256), 0, P->st, LS, P->leaf_start, src, dst,
                       mat, pre, post, P->doffs, c ? P->m2l_list : nullptr, c ? P->m2l_count : nullptr, P->m2l_cap);
}

// (KS, MT) = (ceil(p/4), K/16): the four k-step counts that round up to K
template <int MODE, int K>
static void launch_xlat(eb_fmm_plan P, const XlatLevels& LS, const double2* src, double2* dst,
                        const double2* pre, const double2* post) {
  switch ((P->p + 3) / 4 - (K / 4 - 3)) {
    case 0: launch_xlat_ks<MODE, K, K / 4 - 3>(P, LS, src, dst, pre, post); break;
    case 1: launch_xlat_ks<MODE, K, K / 4 - 2>(P, LS, src, dst, pre, post); break;
    case 2: launch_xlat_ks<MODE, K, K / 4 - 1>(P, LS, src, dst, pre, post); break;
    default: launch_xlat_ks<MODE, K, K / 4>(P, LS, src, dst, pre, post); break;
  }
}

// coarse levels [a, b] of M2M (a > b, finest first) or L2L (a < b) as subtree launches
static XlatSub xlat_sub(eb_fmm_plan P, int root, int lfirst, int llast, int dl) {
  XlatSub X{};
  X.root = root;
  X.L = P->L;
  const int step = lfirst <= llast ? 1 : -1;
  for (int l = lfirst;; l += step) {
    const int e = X.nlev++;
    X.lev[e] = l;
    X.Wt[e] = P->lvl_W[l];
    X.offt[e] = (long long)P->lvl_off[l];
    X.Ws[e] = P->lvl_W[l + dl];
    X.NBPs[e] = P->lvl_NBP[l + dl];
    X.offs[e] = (long long)P->lvl_off[l + dl];
    if (l == llast) break;
  }
  return X;
}

template <int MODE, int K, int KS>
static void launch_xlat_sub_ks(eb_fmm_plan P, const XlatSub& X, const double2* src, double2* dst, const double2* pre,
                               const double2* post) {
  constexpr int TR = (K == 48 && KS == 9) ? 4 : 0;
  constexpr int MT = TR ? 2 : K / 16;
  hipLaunchKernelGGL((k_xlat_sub<MODE, K, KS, MT, TR>), dim3(1u << (2 * X.root)), dim3(256), 0, P->st, X,
                     P->leaf_start, src, dst, P->dMat + (size_t)MODE * K * K, pre, post);
}

template <int MODE, int K>
static void launch_xlat_sub(eb_fmm_plan P, const XlatSub& X, const double2* src, double2* dst, const double2* pre,
                            const double2* post) {
  switch ((P->p + 3) / 4 - (K / 4 - 3)) {
    case 0: launch_xlat_sub_ks<MODE, K, K / 4 - 3>(P, X, src, dst, pre, post); break;
    case 1: launch_xlat_sub_ks<MODE, K, K / 4 - 2>(P, X, src, dst, pre, post); break;
    case 2: launch_xlat_sub_ks<MODE, K, K / 4 - 1>(P, X, src, dst, pre, post); break;
    default: launch_xlat_sub_ks<MODE, K, K / 4>(P, X, src, dst, pre, post); break;
  }
}

// The upward pass: the finest M2M level alone (its targets fill the chip), then the coarser
// ones three levels per launch (subtrees of 21 targets per block)
template <int K>
static void launch_m2m(eb_fmm_plan P) {
  const int L = P->L;
  if (L - 1 >= 2) launch_xlat<XL_M2M, K>(P, xlat_levels(P, L - 1, L - 1, +1), P->mp, P->mp, P->dPM, P->dQM);
  for (int top = L - 2; top >= 2; top -= 3) {
    const int root = std::max(2, top - 2);
    launch_xlat_sub<XL_M2M, K>(P, xlat_sub(P, root, top, root, +1), P->mp, P->mp, P->dPM, P->dQM);
  }
}

// The downward pass into levels 3 .. L-1: the coarse levels two per launch (subtrees of 20
// targets per block), the finest (L - 1) alone
template <int K>
static void launch_l2l(eb_fmm_plan P) {
  const int L = P->L;
  for (int l = 3; l <= L - 2; l += 2) {
    const int last = std::min(l + 1, L - 2);
    launch_xlat_sub<XL_L2L, K>(P, xlat_sub(P, l - 1, l, last, -1), P->lo, P->lo, P->dPL, P->dQL);
  }
  if (L - 1 >= 3) launch_xlat<XL_L2L, K>(P, xlat_levels(P, L - 1, L - 1, -1), P->lo, P->lo, P->dPL, P->dQL);
}

// the output form of a forward (of.kind == OUT_FORWARD) or transposed apply
static OutForm out_form_of(eb_fmm_plan P, bool trans) {
  if (!trans) return OutForm{P->sgn, P->sqb, OUT_FORWARD, nullptr, P->seg_scale, P->segk};
  return OutForm{P->sgn, P->sqb, P->variant == EB_VARIANT_HILBERT ? OUT_T_HILBERT : OUT_T_CADAP, P->wdzs, P->seg_scale,
                 P->segk};
}

// the near field: RSCAL * (the near sum) into out[perm[i]], on stream s
static void launch_p2p(eb_fmm_plan P, hipStream_t s, double2* out, const OutForm& of) {
  const dim3 grid((unsigned)p2p_item_cap(P->n, P->L, P->p2p_tmax));
  if (of.kind == OUT_FORWARD)
    hipLaunchKernelGGL(k_p2p<false>, grid, dim3(64), 0, s, P->L, P->leaf_start, P->p2p_items,
                       P->p2p_off + (1 << (2 * P->L)), P->zs, P->qa, P->qb, (const double2*)nullptr, P->perm, P->geom,
                       of, out);
  else
    hipLaunchKernelGGL(k_p2p<true>, grid, dim3(64), 0, s, P->L, P->leaf_start, P->p2p_items,
                       P->p2p_off + (1 << (2 * P->L)), P->zs, P->qa, P->qb, (const double2*)nullptr, P->perm, P->geom,
                       of, out);
}

// The far-field passes on the plan's stream, then L2P, which adds RSCAL * (the far field) to the
// near field's output: after the join with the near field's stream, or (stage timing) after
// running the near field in place on the plan's stream.
template <int K, bool TRANS>
static void launch_far(eb_fmm_plan P, double2* out, const OutForm& of) {
  hipStream_t st = P->st;
  const int L = P->L;
  if (P->timing) (void)hipEventRecord(P->ev[1], st);
  // P2M at the finest level
  if constexpr (K >= 48)
    hipLaunchKernelGGL(k_p2m_wide<K>, dim3(1 << (2 * L)), dim3(64), 0, st, L, P->p, P->leaf_start, P->zs, P->qa,
                       P->qb, P->mp + P->lvl_off[L], P->lvl_W[L], P->lvl_NBP[L]);
  else
    hipLaunchKernelGGL(k_p2m<K>, dim3(1 << (2 * L)), dim3(64), 0, st, L, P->p, P->leaf_start, P->zs, P->qa, P->qb,
                     P->mp + P->lvl_off[L], P->lvl_W[L], P->lvl_NBP[L]);
  if (P->timing) (void)hipEventRecord(P->ev[2], st);
  // upward pass
  launch_m2m<K>(P);
  if (P->timing) (void)hipEventRecord(P->ev[3], st);
  // interaction lists of every level in one launch, then the downward pass
  launch_xlat<XL_M2L, K>(P, xlat_levels(P, 2, L, 0), P->mp, P->lo, P->dU, P->dNU);
  if (P->timing) (void)hipEventRecord(P->ev[4], st);
  // the downward pass; the shift into the leaves is left to L2P, which evaluates the parents'
  // expansions directly
  launch_l2l<K>(P);
  if (P->timing) {
    (void)hipEventRecord(P->ev[5], st);
    launch_p2p(P, st, out, of);
    (void)hipEventRecord(P->ev[6], st);
  } else {
    (void)hipStreamWaitEvent(st, P->ev_join, 0);
  }
  const dim3 lgrid((unsigned)l2p_item_cap(P->n, L));
  const int* nl2p = P->l2p_total;
  if (L >= 3)
    hipLaunchKernelGGL((k_l2p<K, true, TRANS>), lgrid, dim3(L2P_T), 0, st, L, P->p, P->l2p_items, nl2p, P->keys,
                       P->zs, P->lo + P->lvl_off[L], P->lvl_W[L], P->lo + P->lvl_off[L - 1], P->lvl_W[L - 1], of,
                       P->perm, P->geom, out);
  else
    hipLaunchKernelGGL((k_l2p<K, false, TRANS>), lgrid, dim3(L2P_T), 0, st, L, P->p, P->l2p_items, nl2p, P->keys,
                       P->zs, P->lo + P->lvl_off[L], P->lvl_W[L], P->lo + P->lvl_off[L], P->lvl_W[L], of, P->perm,
                       P->geom, out);
}

template <int K>
static void launch_far_of(eb_fmm_plan P, double2* out, const OutForm& of) {
  if (of.kind == OUT_FORWARD)
    launch_far<K, false>(P, out, of);
  else
    launch_far<K, true>(P, out, of);
}

// eb_fmm_apply (trans = false) and eb_fmm_apply_transpose (trans = true): the same passes, with
// the transposed charges and output form for the transpose
static int fmm_apply(eb_fmm_plan P, const double* dens, double* gradw, void* stream, bool trans) {
  if (!P || !dens || !gradw) {
    eb_set_error(trans ? "eb_fmm_apply_transpose: null argument" : "eb_fmm_apply: null argument");
    return EB_ERR_ARG;
  }
  if (int rc = use_stream(P, stream)) return rc;
  hipStream_t st = P->st;
  const int n = P->n;
  double2* out = (double2*)gradw;
  if (P->timing) {
    for (auto& e : P->ev)
      if (!e) EB_HIP_CHECK(hipEventCreate(&e));
    EB_HIP_CHECK(hipEventRecord(P->ev[0], st));
  }
  const OutForm of = out_form_of(P, trans);
  if (trans)
    hipLaunchKernelGGL(k_charges_t, dim3((n + 255) / 256), dim3(256), 0, st, n, P->perm, (const double2*)dens, of.kind,
                       P->sgn, P->qa, P->qb);
  else
    hipLaunchKernelGGL(k_charges, dim3((n + 255) / 256), dim3(256), 0, st, n, P->perm, (const double2*)dens, P->wdzs,
                       P->sqb, P->variant == EB_VARIANT_HILBERT ? 1 : 0, P->qa, P->qb);
  if (P->mode == EB_MODE_DIRECT) {
    if (P->timing)
      for (int i = 1; i <= 6; ++i) (void)hipEventRecord(P->ev[i], st);
    hipLaunchKernelGGL(k_direct_part, dim3((n + DIR_T - 1) / DIR_T, P->dir_nch), dim3(DIR_T), 0, st, n, P->dir_sch,
                       P->zs, P->qa, P->qb, P->guard, P->dir_part);
    if (trans)
      hipLaunchKernelGGL(k_direct_sum<true>, dim3((n + 255) / 256), dim3(256), 0, st, n, P->dir_nch, P->dir_part,
                         P->perm, P->geom, of, out);
    else
      hipLaunchKernelGGL(k_direct_sum<false>, dim3((n + 255) / 256), dim3(256), 0, st, n, P->dir_nch, P->dir_part,
                         P->perm, P->geom, of, out);
  } else if (P->L >= 2) {
    if (!P->timing) {
      // the near field beside the far-field passes: fork after the charges (measured at 2^20
      // points, p = 36: 1.046 ms per apply against 1.068 with the near field first and 1.063
      // with the fork after P2M)
      EB_HIP_CHECK(hipEventRecord(P->ev_fork, st));
      EB_HIP_CHECK(hipStreamWaitEvent(P->st2, P->ev_fork, 0));
      launch_p2p(P, P->st2, out, of);
      EB_HIP_CHECK(hipEventRecord(P->ev_join, P->st2));
    }
    switch (P->kmax) {
      case 16: launch_far_of<16>(P, out, of); break;
      case 32: launch_far_of<32>(P, out, of); break;
      case 48: launch_far_of<48>(P, out, of); break;
      default: launch_far_of<64>(P, out, of); break;
    }
  } else {
    if (P->timing)
      for (int i = 1; i <= 5; ++i) (void)hipEventRecord(P->ev[i], st);
    launch_p2p(P, st, out, of);
    if (P->timing) (void)hipEventRecord(P->ev[6], st);
  }
  EB_HIP_CHECK(hipGetLastError());
  if (P->timing) {
    EB_HIP_CHECK(hipEventRecord(P->ev[7], st));
    EB_HIP_CHECK(hipEventSynchronize(P->ev[7]));
    float ms[7] = {0};
    for (int i = 0; i < 7; ++i) EB_HIP_CHECK(hipEventElapsedTime(&ms[i], P->ev[i], P->ev[i + 1]));
    // events: charges, P2M, M2M, M2L, L2L, P2P, L2P; reported as charges, P2M, M2M, M2L, L2L, L2P, P2P
    const int order[7] = {0, 1, 2, 3, 4, 6, 5};
    for (int i = 0; i < 7; ++i) P->stage_ms[order[i]] = ms[i];
  }
  return EB_OK;
}

extern "C" int eb_fmm_apply(eb_fmm_plan P, const double* dens, double* gradw, void* stream) {
  return fmm_apply(P, dens, gradw, stream, false);
}

extern "C" int eb_fmm_apply_transpose(eb_fmm_plan P, const double* lam, double* mu, void* stream) {
  return fmm_apply(P, lam, mu, stream, true);
}

extern "C" int eb_fmm_plan_rebuild(eb_fmm_plan P, const double* x, const double* y, const double* wdz,
                                   void* stream) {
  if (!P || !x || !y || !wdz) {
    eb_set_error("eb_fmm_plan_rebuild: null argument");
    return EB_ERR_ARG;
  }
  if (int rc = use_stream(P, stream)) return rc;
  return build_tree(P, x, y, wdz);
}

extern "C" int eb_fmm_set_p2p_tile(eb_fmm_plan P, int32_t tmax) {
  if (!P || tmax < P2P_TMAX_MIN || tmax > P2P_TMAX_MAX) {
    eb_set_error("eb_fmm_set_p2p_tile: tile %d outside %d..%d", tmax, P2P_TMAX_MIN, P2P_TMAX_MAX);
    return EB_ERR_ARG;
  }
  P->p2p_tmax = tmax;
  return P->tree_sorted ? build_items(P, false) : EB_OK;  // an unsorted direct-mode tree builds them when sorted
}

extern "C" int eb_fmm_set_guard(eb_fmm_plan P, const int32_t* flag) {
  if (!P) return (eb_set_error("eb_fmm_set_guard: null plan"), EB_ERR_ARG);
  P->guard = flag;
  return EB_OK;
}

extern "C" int eb_fmm_plan_destroy(eb_fmm_plan P) {
  if (P && P->st) (void)hipStreamSynchronize(P->st);
  plan_free(P);
  return EB_OK;
}

extern "C" int eb_fmm_enable_timing(eb_fmm_plan P, int32_t on) {
  if (!P) return EB_ERR_ARG;
  P->timing = on != 0;
  return EB_OK;
}

extern "C" int eb_fmm_stage_ms(eb_fmm_plan P, double* ms7) {
  if (!P || !ms7) return EB_ERR_ARG;
  for (int i = 0; i < 7; ++i) ms7[i] = P->stage_ms[i];
  return EB_OK;
}

extern "C" int eb_fmm_plan_info(eb_fmm_plan P, double* reals3, int64_t* counts8) {
  if (!P) return EB_ERR_ARG;
  if (int rc = ensure_tree(P)) return rc;
  EB_HIP_CHECK(hipStreamSynchronize(P->st));
  Geom g;
  EB_HIP_CHECK(hipMemcpy(&g, P->geom, sizeof(Geom), hipMemcpyDeviceToHost));
  if (reals3) {
    reals3[0] = g.rscal;
    reals3[1] = g.rminx;
    reals3[2] = g.rminy;
  }
  if (!counts8) return EB_OK;
  const int L = P->L, nside = 1 << L;
  std::vector<int> ls((size_t)(1 << (2 * L)) + 1);
  EB_HIP_CHECK(hipMemcpy(ls.data(), P->leaf_start, sizeof(int) * ls.size(), hipMemcpyDeviceToHost));
  auto cnt = [&](int l, int ix, int iy) -> long long {
    if (ix < 0 || iy < 0 || ix >= (1 << l) || iy >= (1 << l)) return 0;
    uint32_t m = morton((uint32_t)ix, (uint32_t)iy);
    int sh = 2 * (L - l);
    return ls[(size_t)(m + 1) << sh] - ls[(size_t)m << sh];
  };
  long long near = 0, maxleaf = 0, nonempty_leaf = 0;
  for (int iy = 0; iy < nside; ++iy)
    for (int ix = 0; ix < nside; ++ix) {
      long long t = cnt(L, ix, iy);
      if (!t) continue;
      ++nonempty_leaf;
      maxleaf = std::max(maxleaf, t);
      long long s = 0;
      for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) s += cnt(L, ix + dx, iy + dy);
      near += t * s - t;
    }
  long long m2l = 0, boxes = 0;
  if (P->mode == EB_MODE_FMM)
    for (int l = 2; l <= L; ++l) {
      const int ns = 1 << l;
      for (int iy = 0; iy < ns; ++iy)
        for (int ix = 0; ix < ns; ++ix) {
          if (!cnt(l, ix, iy)) continue;
          ++boxes;
          int dxs[27], dys[27];
          int no = intl_offsets(ix & 1, iy & 1, dxs, dys);
          for (int o = 0; o < no; ++o)
            if (cnt(l, ix + dxs[o], iy + dys[o])) ++m2l;
        }
    }
  if (P->mode == EB_MODE_DIRECT) near = (long long)P->n * (P->n - 1);
  counts8[0] = L;
  counts8[1] = P->p;
  counts8[2] = P->kmax;
  counts8[3] = near;
  counts8[4] = m2l;
  counts8[5] = boxes;
  counts8[6] = nonempty_leaf;
  counts8[7] = maxleaf;
  return EB_OK;
}

extern "C" int eb_fmm_plan_m2l_split(eb_fmm_plan P, int64_t* out3) {
  if (!P || !out3) return (eb_set_error("eb_fmm_plan_m2l_split: bad arguments"), EB_ERR_ARG);
  if (int rc = ensure_tree(P)) return rc;
  EB_HIP_CHECK(hipStreamSynchronize(P->st));
  const int L = P->L;
  std::vector<int> ls((size_t)(1 << (2 * L)) + 1);
  EB_HIP_CHECK(hipMemcpy(ls.data(), P->leaf_start, sizeof(int) * ls.size(), hipMemcpyDeviceToHost));
  auto cnt = [&](int l, int ix, int iy) -> long long {
    if (ix < 0 || iy < 0 || ix >= (1 << l) || iy >= (1 << l)) return 0;
    uint32_t m = morton((uint32_t)ix, (uint32_t)iy);
    int sh = 2 * (L - l);
    return ls[(size_t)(m + 1) << sh] - ls[(size_t)m << sh];
  };
  long long grp[2] = {0, 0};
  if (P->mode == EB_MODE_FMM)
    for (int l = 2; l <= L; ++l) {
      const int ns = 1 << l;
      for (int iy = 0; iy < ns; ++iy)
        for (int ix = 0; ix < ns; ++ix) {
          if (!cnt(l, ix, iy)) continue;
          int dxs[27], dys[27];
          const int no = m2l_offsets(ix & 1, iy & 1, dxs, dys);
          for (int o = 0; o < no; ++o)
            if (cnt(l, ix + dxs[o], iy + dys[o])) ++grp[o < M2L_NEAR ? 0 : 1];
        }
    }
  out3[0] = grp[0];
  out3[1] = grp[1];
  out3[2] = m2l_far_terms(P->p, P->kmax);
  return EB_OK;
}

// per-level non-empty box counts (levels 1..levmax) of a plan, host copy
extern "C" int eb_fmm_plan_ntbox(eb_fmm_plan P, int32_t* ntbox) {
  if (!P || !ntbox) return EB_ERR_ARG;
  if (int rc = ensure_tree(P)) return rc;
  const int nleaf = 1 << (2 * P->L);
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, P->st, P->ntbox, EB_MAXLEV + 2);
  hipLaunchKernelGGL(k_ntbox, dim3((nleaf + 255) / 256), dim3(256), 0, P->st, P->leaf_start, P->L, P->ntbox);
  EB_HIP_CHECK(hipGetLastError());
  EB_HIP_CHECK(hipMemcpyAsync(ntbox, P->ntbox, sizeof(int) * (P->L + 1), hipMemcpyDeviceToHost, P->st));
  EB_HIP_CHECK(hipStreamSynchronize(P->st));
  return EB_OK;
}

// ------------------------------------------------------------------------------------------
// reference tables on the GPU (parity interface)
// ------------------------------------------------------------------------------------------
__global__ void k_tab_box(const double* __restrict__ x, const double* __restrict__ y, int n, int l,
                          const Geom* __restrict__ g, int32_t* __restrict__ ibox, int* __restrict__ icnt,
                          int* __restrict__ first, uint32_t* __restrict__ keys, int32_t* __restrict__ vals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double rs = g->rscal;
  const double X = (x[i] - g->rminx) * rs;
  const double Y = (y[i] - g->rminy) * rs;
  const int nside = 1 << l;
  const double h = 64.0 / nside;
  int ix = (int)(X / h), iy = (int)(Y / h);
  ix = ix >= nside ? nside - 1 : ix;
  iy = iy >= nside ? nside - 1 : iy;
  ix = ix < 0 ? 0 : ix;
  iy = iy < 0 ? 0 : iy;
  const int b = iy * nside + ix;  // 0-based row-major (cadap2.f:320-322)
  ibox[i] = b + 1;
  atomicAdd(&icnt[b], 1);
  atomicMin(&first[b], i);
  keys[i] = (uint32_t)b;
  vals[i] = i;
}

__global__ void k_tab_flags(const int32_t* __restrict__ ibox, const int* __restrict__ first, int n,
                            int* __restrict__ flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flag[i] = (first[ibox[i] - 1] == i) ? 1 : 0;
}

__global__ void k_tab_iloc(const int32_t* __restrict__ ibox, const int* __restrict__ flag,
                           const int* __restrict__ rank, int n, int32_t* __restrict__ iloc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (flag[i]) iloc[ibox[i] - 1] = rank[i] + 1;
}

__global__ void k_tab_final(const int32_t* __restrict__ vals, int n, int32_t* __restrict__ iatadr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  iatadr[i] = vals[i] + 1;
}

__global__ void k_max(const int* __restrict__ a, int n, int* __restrict__ out) {
  int m = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = max(m, a[i]);
  atomicMax(out, m);
}

extern "C" int eb_cadap_tables(int32_t ilev, int32_t n, const double* x, const double* y, int32_t* ibox,
                               int32_t* icnt, int32_t* ioffst, int32_t* iatadr, int32_t* iloc, int32_t* ntbox,
                               int32_t* nsmax, void* stream) {
  if (ilev < 1 || ilev > 15 || n < 1) {
    eb_set_error("eb_cadap_tables: ilev=%d n=%d", ilev, n);
    return EB_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  const int nb = 1 << (2 * ilev);
  const int bbblocks = std::min(1024, (n + BB_THREADS - 1) / BB_THREADS);
  Geom* g = nullptr;
  double4* part = nullptr;
  int *first = nullptr, *flag = nullptr, *rank = nullptr, *tmp = nullptr, *dmax = nullptr, *stmp = nullptr;
  uint32_t *k0 = nullptr, *k1 = nullptr;
  int32_t *v0 = nullptr, *v1 = nullptr;
  int rc = EB_OK;
  auto cleanup = [&]() {
    void* ptrs[] = {g, part, first, flag, rank, tmp, dmax, k0, k1, v0, v1, stmp};
    for (void* q : ptrs)
      if (q) (void)hipFree(q);
  };
  if ((rc = dmalloc(&g, 1)) || (rc = dmalloc(&part, bbblocks)) || (rc = dmalloc(&first, nb)) ||
      (rc = dmalloc(&flag, n)) || (rc = dmalloc(&rank, n)) || (rc = dmalloc(&tmp, radix_sort_temp_ints(n))) ||
      (rc = dmalloc(&dmax, 1)) || (rc = dmalloc(&k0, n)) || (rc = dmalloc(&k1, n)) || (rc = dmalloc(&v0, n)) ||
      (rc = dmalloc(&v1, n)) || (rc = dmalloc(&stmp, scan_temp_ints(std::max(n, nb))))) {
    cleanup();
    return rc;
  }
  hipLaunchKernelGGL(k_bbox_part, dim3(bbblocks), dim3(BB_THREADS), 0, st, x, y, n, part);
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(BB_THREADS), 0, st, part, bbblocks, g);
  bool ok = hipMemsetAsync(icnt, 0, sizeof(int) * nb, st) == hipSuccess &&
            hipMemsetAsync(iloc, 0, sizeof(int) * nb, st) == hipSuccess &&
            hipMemsetAsync(first, 0x7f, sizeof(int) * nb, st) == hipSuccess &&
            hipMemsetAsync(dmax, 0, sizeof(int), st) == hipSuccess;
  if (!ok) {
    cleanup();
    eb_set_error("eb_cadap_tables: memset failed");
    return EB_ERR_HIP;
  }
  const int gb = (n + 255) / 256;
  hipLaunchKernelGGL(k_tab_box, dim3(gb), dim3(256), 0, st, x, y, n, ilev, g, ibox, icnt, first, k0, v0);
  if ((rc = scan_exclusive(icnt, ioffst, nb, true, stmp, st)) ||
      (rc = radix_sort_pairs(k0, v0, k1, v1, n, 2 * ilev, tmp, st))) {
    cleanup();
    return rc;
  }
  hipLaunchKernelGGL(k_tab_final, dim3(gb), dim3(256), 0, st, v0, n, iatadr);
  hipLaunchKernelGGL(k_tab_flags, dim3(gb), dim3(256), 0, st, ibox, first, n, flag);
  if ((rc = scan_exclusive(flag, rank, n, false, stmp, st))) {
    cleanup();
    return rc;
  }
  hipLaunchKernelGGL(k_tab_iloc, dim3(gb), dim3(256), 0, st, ibox, flag, rank, n, iloc);
  hipLaunchKernelGGL(k_max, dim3(std::min(1024, (nb + 255) / 256)), dim3(256), 0, st, icnt, nb, dmax);
  int last_rank = 0, last_flag = 0, hmax = 0;
  ok = hipGetLastError() == hipSuccess &&
       hipMemcpyAsync(&last_rank, rank + (n - 1), sizeof(int), hipMemcpyDeviceToHost, st) == hipSuccess &&
       hipMemcpyAsync(&last_flag, flag + (n - 1), sizeof(int), hipMemcpyDeviceToHost, st) == hipSuccess &&
       hipMemcpyAsync(&hmax, dmax, sizeof(int), hipMemcpyDeviceToHost, st) == hipSuccess &&
       hipStreamSynchronize(st) == hipSuccess;
  cleanup();
  if (!ok) {
    eb_set_error("eb_cadap_tables: kernel or copy failed");
    return EB_ERR_HIP;
  }
  if (ntbox) *ntbox = last_rank + last_flag;
  if (nsmax) *nsmax = hmax;
  return EB_OK;
}

// MKINTL / MKNBOR for every box of a level, with the domain clipping of cadap2.f:727-730,801-815
__global__ void k_lists(int l, int32_t* __restrict__ intl, int32_t* __restrict__ nintl, int32_t* __restrict__ nbor,
                        int32_t* __restrict__ nnbor) {
  const int nside = 1 << l;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nside * nside) return;
  const int ix = b % nside, iy = b / nside;
  int dxs[27], dys[27];
  const int no = intl_offsets(ix & 1, iy & 1, dxs, dys);
  int c = 0;
  for (int o = 0; o < no; ++o) {
    const int jx = ix + dxs[o], jy = iy + dys[o];
    if (jx < 0 || jy < 0 || jx >= nside || jy >= nside) continue;
    intl[(size_t)b * 27 + c++] = jy * nside + jx + 1;
  }
  nintl[b] = c;
  c = 0;
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy) {
      const int jx = ix + dx, jy = iy + dy;
      if (jx < 0 || jy < 0 || jx >= nside || jy >= nside) continue;
      nbor[(size_t)b * 9 + c++] = jy * nside + jx + 1;
    }
  nnbor[b] = c;
}

extern "C" int eb_cadap_lists(int32_t ilev, int32_t* intl, int32_t* nintl, int32_t* nbor, int32_t* nnbor,
                              void* stream) {
  if (ilev < 1 || ilev > 12) {
    eb_set_error("eb_cadap_lists: ilev=%d", ilev);
    return EB_ERR_ARG;
  }
  const int nb = 1 << (2 * ilev);
  hipLaunchKernelGGL(k_lists, dim3((nb + 255) / 256), dim3(256), 0, (hipStream_t)stream, ilev, intl, nintl, nbor,
                     nnbor);
  EB_HIP_CHECK(hipGetLastError());
  EB_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  return EB_OK;
}

// ==========================================================================================
// The FMM on the adaptive tree of DAPIF2 (eb_afmm_*): the same kernel sums on the tree that
// DAASSI .. DALI34 build (dapif1.f:1165-1680, dapif2.f:3-474; eb_atree), whose childless boxes
// hold at most NAPB points, in place of the uniform grid of LEVMAX levels.
//
// Passes (Greengard-Rokhlin adaptive form):
//   P2M at every childless box; M2M into every parent, level by level upward;
//   M2L from the 27 slots of MKINTL's interaction list (cadap2.f:780-828: children of the
//     parent's colleagues, not adjacent) where such a box exists, every box of level >= 2;
//   P2L from list X: the points of a childless box c into the local expansion of every box b
//     of c's list W (b is then one of c's X list's targets);
//   L2L into every box of level >= 3, level by level downward;
//   at every childless box b: L2P, the multipole expansions of list W (descendants of b's
//     colleagues that are not adjacent to b but whose parents are), and the direct sum over
//     list U (every childless box adjacent to b, b itself included).
// W and U come from a descent through the colleagues; with the interaction lists of b's
// ancestors they count every source exactly once (eb_afmm_check verifies it on the host).
// The translations run on the matrix-core kernel above (xlat_body, ADAPT): the expansions of
// box k sit at k * 2 KMAX (both series), scaled by the box's half-width, so one table serves
// every level.  Coordinates are the tree's own: (x - XONEW) * RSCAL in [-64, 64]^2.
// ==========================================================================================
struct eb_afmm_s {
  int n = 0, p = 0, kmax = 0, variant = 0, nbox = 0, nleaf = 0, nlev = 0;
  double sgn = -1.0, sqb = -1.0, rscal = 1.0, xonew = 0.0, yonew = 0.0;
  hipStream_t st = nullptr;
  // host copies of the box data and lists (eb_afmm_check, eb_afmm_info)
  std::vector<int> lev, bx, by, par, leafv, bstart, bcnt;
  std::vector<std::vector<int>> U, W, X;
  std::vector<int> stab27_h;
  long long cnt[8] = {0};
  // device
  Geom* geom = nullptr;
  int32_t* perm = nullptr;
  double2 *zs = nullptr, *wdzs = nullptr, *qa = nullptr, *qb = nullptr, *mp = nullptr, *lo = nullptr;
  double2* bctr = nullptr;  // box centres
  double* brad = nullptr;   // box half-widths
  int *dstart = nullptr, *dcnt = nullptr;  // point range of each childless box (sorted order)
  int *leaves = nullptr, *stab27 = nullptr, *kid = nullptr, *dpar = nullptr;
  int4* items = nullptr;  // (childless box, first point, end point), at most AF_T points each
  int nitems = 0;
  int *u_off = nullptr, *u_box = nullptr, *w_off = nullptr, *w_box = nullptr;  // per childless box
  int *x_tgt = nullptr, *x_off = nullptr, *x_box = nullptr;                     // per X target
  int nxt = 0;
  // per-class target lists of the translation launches: [0] M2L, then M2M per level, L2L per level
  struct Group {
    int* list = nullptr;  // 4 x cap
    int* count = nullptr; // 4
    int cap = 0, grid = 0;
  };
  std::vector<Group> m2m, l2l;
  Group m2l;
  double* dMat = nullptr;
  double2 *dU = nullptr, *dNU = nullptr, *dPM = nullptr, *dQM = nullptr, *dPL = nullptr, *dQL = nullptr;
  std::vector<void*> owned;  // every device allocation above
  eb_atree tree = nullptr;   // held until the points are gathered
};

namespace {

constexpr int AF_T = 64;     // points per childless-box work item
constexpr int AF_CH = 128;   // sources staged per chunk in the near field
constexpr int AF_WG = 4;     // list-W multipoles staged per round in k_af_far

// P2M of every childless box: one wave per box
template <int KMAX>
__global__ __launch_bounds__(64) void k_af_p2m(int p, const int* __restrict__ leaves, const int* __restrict__ start,
                                               const int* __restrict__ cnt, const double2* __restrict__ ctr,
                                               const double* __restrict__ rad, const double2* __restrict__ zs,
                                               const double2* __restrict__ qa, const double2* __restrict__ qb,
                                               double2* __restrict__ mp) {
  const int b = leaves[blockIdx.x];
  const int s = start[b];
  p2m_box<KMAX>(p, s, s + cnt[b], ctr[b], 1.0 / rad[b], zs, qa, qb, mp + (size_t)b * 2 * KMAX);
}

// P2L of list X: the points of every childless box c in X(t) into the local expansion of t
// (centre ct, half-width r), added to what M2L wrote.  With s = r / (z_j - ct), point j adds
//   phi: (qa_j / r) s^(m+1),   psi: (qb_j / r) s^(m+1) + (m + 1) (qc_j / r^2) s^(m+2)
// to coefficient m (qc = -conj(z) qa), the expansion of qa/(z_j - z) and of
// qb/(z_j - z) + qc/(z_j - z)^2 in powers of (z - ct) / r.  One wave per target box, lane m
// owns coefficient m of both series.
template <int KMAX>
__global__ __launch_bounds__(64) void k_af_p2l(int p, const int* __restrict__ tgt, const int* __restrict__ off,
                                               const int* __restrict__ src, const int* __restrict__ start,
                                               const int* __restrict__ cnt, const double2* __restrict__ ctr,
                                               const double* __restrict__ rad, const double2* __restrict__ zs,
                                               const double2* __restrict__ qa, const double2* __restrict__ qb,
                                               double2* __restrict__ lo) {
  constexpr int CH = 8, QK = KMAX / 8 + 1;  // powers s^1 .. s^(KMAX+1): indices 0 .. KMAX
  __shared__ double2 P[CH][KMAX + 8];
  __shared__ double2 AB[CH][3];
  const int t = tgt[blockIdx.x];
  const double2 c = ctr[t];
  const double r = rad[t], ir = 1.0 / r;
  const int lane = threadIdx.x;
  const int js = lane & (CH - 1), qk = lane / CH;
  const int k = lane < KMAX ? lane : KMAX - 1;
  double2 a0 = make_double2(0, 0), a1 = make_double2(0, 0), a2 = make_double2(0, 0);
  for (int q = off[blockIdx.x]; q < off[blockIdx.x + 1]; ++q) {
    const int sb = src[q], s0 = start[sb], e0 = s0 + cnt[sb];
    for (int c0 = s0; c0 < e0; c0 += CH) {
      const int nj = min(CH, e0 - c0);
      if (js < nj) {
        const int j = c0 + js;
        const double2 z = zs[j];
        const double2 sv = cscale(crcp(make_double2(z.x - c.x, z.y - c.y)), r);  // r / (z - ct)
        if (qk == 0) {
          const double2 q1 = qa[j], q2 = qb[j];
          AB[js][0] = cscale(q1, ir);
          AB[js][1] = cscale(q2, ir);
          AB[js][2] = cscale(cmul(make_double2(-z.x, z.y), q1), ir * ir);  // qc / r^2
        }
        // lane qk writes s^(1 + qk + 8 m) at index qk + 8 m
        const double2 s2 = cmul(sv, sv), s4 = cmul(s2, s2), s8 = cmul(s4, s4);
        double2 pw = sv;
        if (qk & 1) pw = cmul(pw, sv);
        if (qk & 2) pw = cmul(pw, s2);
        if (qk & 4) pw = cmul(pw, s4);
#pragma unroll
        for (int m = 0; m < QK; ++m) {
          const int idx = qk + 8 * m;
          if (idx <= KMAX) P[js][idx] = pw;
          if (m + 1 < QK) pw = cmul(pw, s8);
        }
      }
      __syncthreads();
      for (int jj = 0; jj < nj; ++jj) {
        const double2 pk = P[jj][k], pk1 = P[jj][k + 1];
        cfma(a0, AB[jj][0], pk);
        cfma(a1, AB[jj][1], pk);
        cfma(a2, AB[jj][2], pk1);
      }
      __syncthreads();
    }
  }
  rfma(a1, (double)(lane + 1), a2);
  if (lane < p) {
    double2* box = lo + (size_t)t * 2 * KMAX;
    box[lane] = cadd(box[lane], a0);
    box[KMAX + lane] = cadd(box[KMAX + lane], a1);
  }
}

// At the points of one work item of a childless box b: its local expansion (L2P) and the
// multipole expansions of list W (M2P: phi = sum_k a_k w^(k+1), w = r_c / (z - c_c), phi' =
// -(w^2 / r_c) d/dw, psi likewise), combined as GW = phi + sgn (z conj(phi') + conj(psi)) and
// added, times RSCAL, to the near field already in out[perm[i]].
template <int KMAX>
__global__ __launch_bounds__(AF_T) void k_af_far(int p, const int4* __restrict__ items,
                                                 const int* __restrict__ w_off, const int* __restrict__ w_box,
                                                 const double2* __restrict__ ctr, const double* __restrict__ rad,
                                                 const double2* __restrict__ zs, const double2* __restrict__ lo,
                                                 const double2* __restrict__ mp, double sgn,
                                                 const int32_t* __restrict__ perm, const Geom* __restrict__ g,
                                                 double2* __restrict__ out) {
  __shared__ double2 C[2][KMAX];
  __shared__ double2 CW[AF_WG][2][KMAX];
  __shared__ double2 WC[AF_WG];
  __shared__ double WR[AF_WG];
  const int4 it = items[blockIdx.x];
  const int b = it.x;
  const int i = it.y + (int)threadIdx.x;
  const bool act = i < it.z;
  const double2 z = act ? zs[i] : make_double2(0, 0);
  for (int t = threadIdx.x; t < 2 * p; t += AF_T) {
    const int sr = t / p, m = t - sr * p;
    C[sr][m] = lo[(size_t)b * 2 * KMAX + sr * KMAX + m];
  }
  __syncthreads();
  double2 phi = make_double2(0, 0), dphi = make_double2(0, 0), psi = make_double2(0, 0);
  {
    const double2 cb = ctr[b];
    const double ir = 1.0 / rad[b];
    const double2 u = make_double2((z.x - cb.x) * ir, (z.y - cb.y) * ir);
#pragma unroll 4
    for (int m = p - 1; m >= 0; --m) {
      dphi = cmuladd(dphi, u, phi);
      phi = cmuladd(phi, u, C[0][m]);
      psi = cmuladd(psi, u, C[1][m]);
    }
    dphi = cscale(dphi, ir);
  }
  // list W, AF_WG boxes' multipoles staged per round
  const int w0 = w_off[b], w1 = w_off[b + 1];
  for (int q0 = w0; q0 < w1; q0 += AF_WG) {
    const int nq = min(AF_WG, w1 - q0);
    __syncthreads();
    for (int t = threadIdx.x; t < nq * 2 * p; t += AF_T) {
      const int gq = t / (2 * p), r = t - gq * 2 * p, sr = r / p, m = r - sr * p;
      CW[gq][sr][m] = mp[(size_t)w_box[q0 + gq] * 2 * KMAX + sr * KMAX + m];
    }
    if ((int)threadIdx.x < nq) {
      const int c = w_box[q0 + threadIdx.x];
      WC[threadIdx.x] = make_double2(ctr[c].x, ctr[c].y);
      WR[threadIdx.x] = rad[c];
    }
    __syncthreads();
    for (int gq = 0; gq < nq; ++gq) {
      const double2 cc = WC[gq];
      const double rc = WR[gq];
      const double2 w = cscale(crcp(make_double2(z.x - cc.x, z.y - cc.y)), rc);
      // R(w) = sum_k a_k w^k, R'(w), S(w) = sum_k b_k w^k: phi = w R, dphi/dw = R + w R'
      double2 R = make_double2(0, 0), dR = make_double2(0, 0), S = make_double2(0, 0);
#pragma unroll 4
      for (int m = p - 1; m >= 0; --m) {
        dR = cmuladd(dR, w, R);
        R = cmuladd(R, w, CW[gq][0][m]);
        S = cmuladd(S, w, CW[gq][1][m]);
      }
      const double2 w2 = cmul(w, w);
      phi = cmuladd(w, R, phi);
      psi = cmuladd(w, S, psi);
      // phi'(z) = (R + w R') dw/dz, dw/dz = -w^2 / r_c
      const double2 dq = cmuladd(w, dR, R);
      dphi = cmuladd(cscale(w2, -1.0 / rc), dq, dphi);
    }
  }
  if (!act) return;
  const double2 zc = cmul(z, cconj(dphi));
  const double2 t = make_double2(zc.x + psi.x, zc.y - psi.y);
  const double rs = g->rscal;
  const int o = perm[i];
  const double2 nearv = out[o];
  out[o] = make_double2(nearv.x + (phi.x + sgn * t.x) * rs, nearv.y + (phi.y + sgn * t.y) * rs);
}

// The near field of one work item of a childless box b: the sources of every box of U(b)
// (b itself first), staged through LDS per box and chunk; with T < 64 targets the lanes split
// into G = 64 / T slices of the sources, summed through LDS at the end.  Writes RSCAL * (the
// sum) to out[perm[i]].
__global__ __launch_bounds__(64) void k_af_p2p(const int4* __restrict__ items, const int* __restrict__ u_off,
                                               const int* __restrict__ u_box, const int* __restrict__ start,
                                               const int* __restrict__ cnt, const double2* __restrict__ zs,
                                               const double2* __restrict__ qa, const double2* __restrict__ qb,
                                               const int32_t* __restrict__ perm, const Geom* __restrict__ g,
                                               double sgn, double2* __restrict__ out) {
  __shared__ double2 sz[AF_CH], sa[AF_CH], sq[AF_CH];
  const int4 it = items[blockIdx.x];
  const int b = it.x, T = it.z - it.y;
  const int lane = threadIdx.x;
  const int G = T < 64 ? 64 / T : 1;
  const int tl = lane % T, gs = lane / T;
  const bool act = gs < G;
  const int i = it.y + tl;
  const double2 zi = act ? zs[i] : make_double2(0, 0);
  double2 a1 = make_double2(0, 0), a2 = make_double2(0, 0);
  for (int q = u_off[b]; q < u_off[b + 1]; ++q) {
    const int ub = u_box[q], s0 = start[ub], n0 = cnt[ub];
    const bool self = ub == b;
    for (int c0 = 0; c0 < n0; c0 += AF_CH) {
      const int nc = min(AF_CH, n0 - c0);
      __syncthreads();
      for (int f = lane; f < nc; f += 64) {
        const int j = s0 + c0 + f;
        sz[f] = zs[j];
        sa[f] = qa[j];
        sq[f] = qb[j];
      }
      __syncthreads();
      if (act) {
        if (self) {
          const int fi = i - s0 - c0;  // the target's own position in this chunk (may be outside)
          for (int f = gs; f < nc; f += G) p2p_pair<true, true>(zi, sz[f], sa[f], sq[f], f == fi, a1, a2);
        } else {
          for (int f = gs; f < nc; f += G) p2p_pair<false, true>(zi, sz[f], sa[f], sq[f], false, a1, a2);
        }
      }
    }
  }
  __syncthreads();
  sz[lane] = make_double2(a1.x + sgn * a2.x, a1.y + sgn * a2.y);
  __syncthreads();
  if (lane < T) {
    double2 s = make_double2(0, 0);
    for (int q = 0; q < G; ++q) s = cadd(s, sz[lane + q * T]);
    const double rs = g->rscal;
    out[perm[it.y + lane]] = make_double2(s.x * rs, s.y * rs);
  }
}

// sorted coordinates and weights: zs[i] = (xs, ys)[perm[i]], wdzs[i] = wdz[perm[i]]
__global__ void k_af_gather(int n, const int32_t* __restrict__ perm, const double* __restrict__ xs,
                            const double* __restrict__ ys, const double2* __restrict__ wdz, double2* __restrict__ zs,
                            double2* __restrict__ wdzs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = perm[i];
  zs[i] = make_double2(xs[j], ys[j]);
  wdzs[i] = wdz[j];
}

__global__ void k_af_geom(const double* __restrict__ tgeom, Geom* __restrict__ g) {
  if (threadIdx.x == 0) *g = Geom{tgeom[0], 0.0, 0.0, 0.0};
}

}  // namespace

static void afmm_free(eb_afmm P) {
  if (!P) return;
  if (P->st) (void)hipStreamSynchronize(P->st);
  if (P->tree) eb_atree_destroy(P->tree);
  for (void* q : P->owned)
    if (q) (void)hipFree(q);
  delete P;
}

template <typename T>
static int af_alloc(eb_afmm P, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  const hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
  if (e != hipSuccess) {
    eb_set_error("eb_afmm_create: hipMalloc(%zu bytes) failed: %s", sizeof(T) * count, hipGetErrorString(e));
    *p = nullptr;
    return EB_ERR_NOMEM;
  }
  P->owned.push_back((void*)*p);
  return EB_OK;
}

template <typename T>
static int af_upload(eb_afmm P, T** p, const std::vector<T>& v) {
  int rc = af_alloc(P, p, v.size());
  if (rc != EB_OK) return rc;
  if (!v.empty()) EB_HIP_CHECK(hipMemcpyAsync(*p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, P->st));
  return EB_OK;
}

// Per-class target lists of the translation kernel (ADAPT).  The kernel runs the four
// classes' blocks of one list position side by side, so aligned lists share their sources in
// L2: by_parent puts child c of the i-th parent at position i of class c's list (-1 where
// the child does not exist); otherwise each class lists its boxes in order.
static int af_group(eb_afmm P, const std::vector<int>& boxes, eb_afmm_s::Group& G, bool by_parent = false,
                    const std::vector<std::array<int, 4>>* kids = nullptr) {
  std::vector<int> per[4];
  if (by_parent) {
    for (int q : boxes)
      for (int c = 0; c < 4; ++c) per[c].push_back((*kids)[q][c]);
  } else {
    for (int b : boxes) per[(P->bx[b] & 1) | ((P->by[b] & 1) << 1)].push_back(b);
  }
  int cap = 0, mx = 0;
  for (auto& v : per) cap = std::max(cap, (int)v.size());
  cap = std::max(cap, 1);
  std::vector<int> list((size_t)4 * cap, -1), count(4);
  for (int c = 0; c < 4; ++c) {
    std::copy(per[c].begin(), per[c].end(), list.begin() + (size_t)c * cap);
    count[c] = (int)per[c].size();
    mx = std::max(mx, (count[c] + 15) / 16);
  }
  G.cap = cap;
  G.grid = (4 * mx + 7) / 8 * 8;  // the kernel's own count of logical blocks (xlat_body)
  int rc;
  if ((rc = af_upload(P, &G.list, list)) != EB_OK || (rc = af_upload(P, &G.count, count)) != EB_OK) return rc;
  return EB_OK;
}

// U and W of every childless box by a descent through the colleagues, X from W, and the M2L
// slot table; counts for eb_afmm_info
static void af_lists(eb_afmm P) {
  const int nb = P->nbox;
  auto key = [](int l, int ix, int iy) -> uint64_t {
    return ((uint64_t)l << 58) | ((uint64_t)(uint32_t)ix << 29) | (uint64_t)(uint32_t)iy;
  };
  std::vector<std::pair<uint64_t, int>> idx(nb);
  for (int b = 0; b < nb; ++b) idx[b] = {key(P->lev[b], P->bx[b], P->by[b]), b};
  std::sort(idx.begin(), idx.end());
  auto find = [&](int l, int ix, int iy) -> int {
    if (ix < 0 || iy < 0 || ix >= (1 << l) || iy >= (1 << l)) return -1;
    const uint64_t k = key(l, ix, iy);
    auto it = std::lower_bound(idx.begin(), idx.end(), std::make_pair(k, -1));
    return (it != idx.end() && it->first == k) ? it->second : -1;
  };
  // closed boxes touch or overlap, compared at the finer level
  auto adjacent = [&](int a, int b) {
    const int f = std::max(P->lev[a], P->lev[b]);
    const long sa = 1L << (f - P->lev[a]), sb = 1L << (f - P->lev[b]);
    const long ax0 = P->bx[a] * sa, ax1 = ax0 + sa, ay0 = P->by[a] * sa, ay1 = ay0 + sa;
    const long bx0 = P->bx[b] * sb, bx1 = bx0 + sb, by0 = P->by[b] * sb, by1 = by0 + sb;
    return ax0 <= bx1 && bx0 <= ax1 && ay0 <= by1 && by0 <= ay1;
  };
  std::vector<std::array<int, 4>> kids(nb);
  for (auto& k : kids) k = {-1, -1, -1, -1};
  for (int b = 0; b < nb; ++b)
    if (P->par[b] >= 0) kids[P->par[b]][(P->bx[b] & 1) | ((P->by[b] & 1) << 1)] = b;
  P->U.assign(nb, {});
  P->W.assign(nb, {});
  P->X.assign(nb, {});
  std::vector<int> stack;
  for (int b = 0; b < nb; ++b) {
    if (!P->leafv[b]) continue;
    auto& U = P->U[b];
    auto& W = P->W[b];
    U.push_back(b);
    const int lb = P->lev[b];
    // childless boxes of the same or a coarser level that touch b
    for (int l = 1; l <= lb; ++l) {
      const int ax = P->bx[b] >> (lb - l), ay = P->by[b] >> (lb - l);
      for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy) {
          const int c = find(l, ax + dx, ay + dy);
          if (c < 0 || c == b || !P->leafv[c]) continue;
          if (l < lb && dx == 0 && dy == 0) continue;  // b's own ancestor
          if (adjacent(b, c)) U.push_back(c);
        }
    }
    // finer boxes: the descendants of b's colleagues
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy) {
        if (dx == 0 && dy == 0) continue;
        const int c = find(lb, P->bx[b] + dx, P->by[b] + dy);
        if (c < 0 || P->leafv[c]) continue;
        stack.assign(1, c);
        while (!stack.empty()) {
          const int a = stack.back();
          stack.pop_back();
          for (int ch : kids[a]) {
            if (ch < 0) continue;
            if (!adjacent(b, ch)) {
              W.push_back(ch);
            } else if (P->leafv[ch]) {
              U.push_back(ch);
            } else {
              stack.push_back(ch);
            }
          }
        }
      }
  }
  // Where a childless box of W(b) holds few points, its points act on b's directly (U) instead
  // of through its multipole; where b itself holds few points, the reverse interaction (b in
  // X(c): b's points into c's local expansion) becomes direct too when c is childless.  A
  // direct pair costs 50 flops, an M2P evaluation about 30 p, a P2L term about 26 p per
  // source point plus its share of c's L2P.
  constexpr int W_DIRECT = 24, X_DIRECT = 32;
  for (int b = 0; b < nb; ++b) {
    if (!P->leafv[b]) continue;
    std::vector<int> keep;
    for (int c : P->W[b]) {
      if (P->leafv[c] && P->bcnt[c] <= W_DIRECT) {
        P->U[b].push_back(c);
      } else {
        keep.push_back(c);
      }
      // the reverse: b's points into c (X(c) gets b) unless c is a small childless box
      if (P->leafv[c] && P->bcnt[c] <= X_DIRECT)
        P->U[c].push_back(b);
      else
        P->X[c].push_back(b);
    }
    P->W[b].swap(keep);
  }
  // M2L slots
  P->stab27_h.assign((size_t)nb * 27, nb);
  long long m2l = 0;
  for (int b = 0; b < nb; ++b) {
    if (P->lev[b] < 2) continue;
    int dxs[27], dys[27];
    const int no = m2l_offsets(P->bx[b] & 1, P->by[b] & 1, dxs, dys);
    for (int o = 0; o < no; ++o) {
      const int c = find(P->lev[b], P->bx[b] + dxs[o], P->by[b] + dys[o]);
      if (c >= 0) {
        P->stab27_h[(size_t)b * 27 + o] = c;
        ++m2l;
      }
    }
  }
  long long near = 0, m2p = 0, p2l = 0, xl = 0;
  for (int b = 0; b < nb; ++b) {
    if (P->leafv[b]) {
      for (int c : P->U[b]) near += (long long)P->bcnt[b] * P->bcnt[c];
      m2p += (long long)P->bcnt[b] * P->W[b].size();
    }
    for (int c : P->X[b]) p2l += P->bcnt[c];
    xl += P->X[b].size();
  }
  long long shifted = 0;
  for (int b = 0; b < nb; ++b) shifted += P->lev[b] >= 3;
  P->cnt[0] = nb;
  P->cnt[1] = P->nleaf;
  P->cnt[2] = P->nlev;
  P->cnt[3] = near - P->n;  // ordered pairs i != j
  P->cnt[4] = m2l;
  P->cnt[5] = m2p;
  P->cnt[6] = p2l;
  P->cnt[7] = shifted;  // M2M and L2L translations each
  (void)xl;
}

extern "C" int eb_afmm_create(eb_afmm* out, int32_t n, const double* x, const double* y, const double* wdz,
                              int32_t napb, int32_t nterms, int32_t variant, void* stream) {
  if (!out || !x || !y || !wdz || n < 2 || napb < 1 || nterms < 4 || nterms > 60 || variant < 0 ||
      variant > EB_VARIANT_HILBERT) {
    eb_set_error("eb_afmm_create: bad arguments (n >= 2, napb >= 1, 4 <= nterms <= 60)");
    return EB_ERR_ARG;
  }
  *out = nullptr;
  eb_afmm P = new eb_afmm_s();
  P->n = n;
  P->p = nterms;
  P->kmax = pick_kmax(nterms);
  P->variant = variant;
  P->st = (hipStream_t)stream;
  P->sgn = (variant == EB_VARIANT_DR) ? -1.0 : (variant == EB_VARIANT_HILBERT ? 0.0 : 1.0);
  P->sqb = (variant == EB_VARIANT_PLUS) ? 1.0 : -1.0;
  const int K = P->kmax;
#define AF_TRY(x)           \
  do {                      \
    int _rc = (x);          \
    if (_rc != EB_OK) {     \
      afmm_free(P);         \
      return _rc;           \
    }                       \
  } while (0)
  // the tree (its build synchronises the stream per level)
  int32_t ierr[3] = {0, 0, 0};
  // at most 29 levels: box positions then fit the 29-bit fields of the host's box keys
  int rc = eb_atree_build(&P->tree, n, x, y, napb, 29, 0.0, ierr, stream);
  eb_atree T = P->tree;
  if (rc != EB_OK) {
    if (rc == EB_ATREE_TOO_DEEP)
      eb_set_error("eb_afmm_create: the tree needs more than 29 levels (more than %d coincident points?)", napb);
    else if (rc == EB_ATREE_FEW_POINTS)
      eb_set_error("eb_afmm_create: %d points do not exceed NAPB = %d", n, napb);
    afmm_free(P);
    return rc < EB_ERR_UNSUPPORTED ? EB_ERR_ARG : rc;
  }
  int32_t c8[8];
  eb_atree_view V;
  eb_atree_counts(T, c8);
  eb_atree_tables(T, &V);
  const int nb = c8[0], nlev = c8[1], jatlen = c8[6];
  P->nbox = nb;
  P->nlev = nlev;
  std::vector<double> xb(nb), yb(nb), tg(3);
  std::vector<int32_t> nl(nlev + 2), npar(nb), indb(nb), iat(nb), nbat(nb), jat(jatlen);
  bool ok = hipMemcpyAsync(xb.data(), V.xb, sizeof(double) * nb, hipMemcpyDeviceToHost, P->st) == hipSuccess &&
            hipMemcpyAsync(yb.data(), V.yb, sizeof(double) * nb, hipMemcpyDeviceToHost, P->st) == hipSuccess &&
            hipMemcpyAsync(tg.data(), V.geom, sizeof(double) * 3, hipMemcpyDeviceToHost, P->st) == hipSuccess &&
            hipMemcpyAsync(nl.data(), V.nl, sizeof(int32_t) * (nlev + 2), hipMemcpyDeviceToHost, P->st) == hipSuccess &&
            hipMemcpyAsync(npar.data(), V.npar, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, P->st) == hipSuccess &&
            hipMemcpyAsync(indb.data(), V.indb, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, P->st) == hipSuccess &&
            hipMemcpyAsync(iat.data(), V.iat, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, P->st) == hipSuccess &&
            hipMemcpyAsync(nbat.data(), V.nbat, sizeof(int32_t) * nb, hipMemcpyDeviceToHost, P->st) == hipSuccess &&
            hipMemcpyAsync(jat.data(), V.jat, sizeof(int32_t) * jatlen, hipMemcpyDeviceToHost, P->st) == hipSuccess &&
            hipStreamSynchronize(P->st) == hipSuccess;
  if (!ok) {
    afmm_free(P);
    eb_set_error("eb_afmm_create: copying the tree failed");
    return EB_ERR_HIP;
  }
  P->rscal = tg[0];
  P->xonew = tg[1];
  P->yonew = tg[2];
  // box levels, integer positions, parents, childless flags; points in childless-box order
  P->lev.assign(nb, 0);
  P->bx.assign(nb, 0);
  P->by.assign(nb, 0);
  P->par.assign(nb, -1);
  P->leafv.assign(nb, 0);
  P->bstart.assign(nb, 0);
  P->bcnt.assign(nb, 0);
  for (int l = 1; l <= nlev; ++l)
    for (int b = nl[l - 1] - 1; b < nl[l] - 1 && b < nb; ++b) P->lev[b] = l;
  std::vector<int32_t> perm;
  perm.reserve(n);
  std::vector<int> leaves;
  for (int b = 0; b < nb; ++b) {
    const int l = P->lev[b];
    const double h = 128.0 / (double)(1 << l);
    P->bx[b] = (int)lround((xb[b] + 64.0) / h - 0.5);
    P->by[b] = (int)lround((yb[b] + 64.0) / h - 0.5);
    P->par[b] = npar[b] - 1;
    if (indb[b] == 1) {
      P->leafv[b] = 1;
      P->bstart[b] = (int)perm.size();
      P->bcnt[b] = nbat[b];
      for (int q = 0; q < nbat[b]; ++q) perm.push_back(jat[iat[b] + q] - 1);  // JAT(IAT(b)+1..): 1-based
      leaves.push_back(b);
    }
  }
  P->nleaf = (int)leaves.size();
  if ((int)perm.size() != n) {
    afmm_free(P);
    eb_set_error("eb_afmm_create: the childless boxes hold %zu of %d points", perm.size(), n);
    return EB_ERR_ARG;
  }
  af_lists(P);
  if ((size_t)2 * K * 16 * (size_t)(nb + 1) >= ((size_t)1 << 31)) {
    afmm_free(P);
    eb_set_error("eb_afmm_create: %d boxes exceed the translation kernel's 2 GiB window", nb);
    return EB_ERR_ARG;
  }
  // device tables
  std::vector<double2> ctr(nb);
  std::vector<double> rad(nb);
  std::vector<int> kid((size_t)nb * 4, nb), par1(nb, nb);
  for (int b = 0; b < nb; ++b) {
    ctr[b] = make_double2(xb[b], yb[b]);
    rad[b] = 64.0 / (double)(1 << P->lev[b]);
    if (P->par[b] >= 0) {
      kid[(size_t)P->par[b] * 4 + ((P->bx[b] & 1) | ((P->by[b] & 1) << 1))] = b;
      par1[b] = P->par[b];
    }
  }
  std::vector<int4> items;
  std::vector<int> uo(nb + 1, 0), ub, wo(nb + 1, 0), wb, xt, xo(1, 0), xbx;
  for (int b = 0; b < nb; ++b) {
    if (P->leafv[b]) {
      const int s = P->bstart[b], c = P->bcnt[b];
      for (int q = 0; q < c; q += AF_T) items.push_back(make_int4(b, s + q, s + std::min(c, q + AF_T), 0));
    }
    ub.insert(ub.end(), P->U[b].begin(), P->U[b].end());
    uo[b + 1] = (int)ub.size();
    wb.insert(wb.end(), P->W[b].begin(), P->W[b].end());
    wo[b + 1] = (int)wb.size();
    if (!P->X[b].empty()) {
      xt.push_back(b);
      xbx.insert(xbx.end(), P->X[b].begin(), P->X[b].end());
      xo.push_back((int)xbx.size());
    }
  }
  P->nitems = (int)items.size();
  P->nxt = (int)xt.size();
  Tables TB;
  build_tables(P->p, K, TB);
  AF_TRY(af_alloc(P, &P->geom, 1));
  AF_TRY(af_upload(P, &P->perm, perm));
  AF_TRY(af_alloc(P, &P->zs, n));
  AF_TRY(af_alloc(P, &P->wdzs, n));
  AF_TRY(af_alloc(P, &P->qa, n));
  AF_TRY(af_alloc(P, &P->qb, n));
  AF_TRY(af_alloc(P, &P->mp, (size_t)nb * 2 * K));
  AF_TRY(af_alloc(P, &P->lo, (size_t)nb * 2 * K));
  // level-1 boxes receive no translation: their local expansions stay zero
  EB_HIP_CHECK(hipMemsetAsync(P->lo, 0, sizeof(double2) * (size_t)nb * 2 * K, P->st));
  AF_TRY(af_upload(P, &P->bctr, ctr));
  AF_TRY(af_upload(P, &P->brad, rad));
  AF_TRY(af_upload(P, &P->dstart, P->bstart));
  AF_TRY(af_upload(P, &P->dcnt, P->bcnt));
  AF_TRY(af_upload(P, &P->leaves, leaves));
  AF_TRY(af_upload(P, &P->stab27, P->stab27_h));
  AF_TRY(af_upload(P, &P->kid, kid));
  AF_TRY(af_upload(P, &P->dpar, par1));
  AF_TRY(af_upload(P, &P->items, items));
  AF_TRY(af_upload(P, &P->u_off, uo));
  AF_TRY(af_upload(P, &P->u_box, ub));
  AF_TRY(af_upload(P, &P->w_off, wo));
  AF_TRY(af_upload(P, &P->w_box, wb));
  AF_TRY(af_upload(P, &P->x_tgt, xt));
  AF_TRY(af_upload(P, &P->x_off, xo));
  AF_TRY(af_upload(P, &P->x_box, xbx));
  AF_TRY(af_upload(P, &P->dMat, TB.Mat));
  AF_TRY(af_upload(P, &P->dU, TB.U));
  AF_TRY(af_upload(P, &P->dNU, TB.NU));
  AF_TRY(af_upload(P, &P->dPM, TB.PM));
  AF_TRY(af_upload(P, &P->dQM, TB.QM));
  AF_TRY(af_upload(P, &P->dPL, TB.PL));
  AF_TRY(af_upload(P, &P->dQL, TB.QL));
  // translation targets: M2L at every box of level >= 2; M2M into the parents of each level
  // (finest first); L2L into every box of each level >= 3 (coarsest first)
  {
    std::vector<std::array<int, 4>> kids(nb);
    for (auto& k : kids) k = {-1, -1, -1, -1};
    for (int b = 0; b < nb; ++b)
      if (P->par[b] >= 0) kids[P->par[b]][(P->bx[b] & 1) | ((P->by[b] & 1) << 1)] = b;
    // M2L into every box of level >= 2: the children of every parent, aligned by parent
    std::vector<int> parents;
    for (int b = 0; b < nb; ++b)
      if (!P->leafv[b]) parents.push_back(b);
    AF_TRY(af_group(P, parents, P->m2l, true, &kids));
    for (int l = nlev - 1; l >= 2; --l) {
      std::vector<int> t;
      for (int b = nl[l - 1] - 1; b < nl[l] - 1; ++b)
        if (!P->leafv[b]) t.push_back(b);
      if (t.empty()) continue;
      P->m2m.emplace_back();
      AF_TRY(af_group(P, t, P->m2m.back()));
    }
    for (int l = 3; l <= nlev; ++l) {
      std::vector<int> t;  // the parents (level l - 1) of the targets
      for (int b = nl[l - 2] - 1; b < nl[l - 1] - 1; ++b)
        if (!P->leafv[b]) t.push_back(b);
      if (t.empty()) continue;
      P->l2l.emplace_back();
      AF_TRY(af_group(P, t, P->l2l.back(), true, &kids));
    }
  }
  // the points in childless-box order, in the tree's scaled coordinates (x - XONEW) * RSCAL
  hipLaunchKernelGGL(k_af_geom, dim3(1), dim3(64), 0, P->st, V.geom, P->geom);
  hipLaunchKernelGGL(k_af_gather, dim3((n + 255) / 256), dim3(256), 0, P->st, n, P->perm, V.xs, V.ys,
                     (const double2*)wdz, P->zs, P->wdzs);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(P->st) != hipSuccess) {
    afmm_free(P);
    eb_set_error("eb_afmm_create: setup kernels failed");
    return EB_ERR_HIP;
  }
  eb_atree_destroy(P->tree);
  P->tree = nullptr;
#undef AF_TRY
  *out = P;
  return EB_OK;
}

template <int MODE, int K, int KS>
static void af_xlat_ks(eb_afmm P, const eb_afmm_s::Group& G, const int* stab, const double2* src, double2* dst,
                       const double2* pre, const double2* post) {
  constexpr int TR = (K == 48 && KS == 9) ? 4 : 0;  // as launch_xlat_ks
  constexpr int MT = TR ? 2 : K / 16;
  XlatLevels LS{};
  LS.nent = 1;
  LS.total = G.grid;
  const double* mat = P->dMat + (size_t)MODE * K * K;
  // no register cap: the uniform M2L under its two-waves cap holds 246 of 256 VGPRs, and the
  // slot-table addressing pushed this one over it (404 B per lane of scratch, 46 times the
  // uniform M2L's HBM reads)
  hipLaunchKernelGGL((k_axlat<MODE, K, KS, MT, TR>), dim3(G.grid), dim3(256), 0, P->st, LS, (const int*)nullptr, src,
                     dst, mat, pre, post, (const int*)nullptr, G.list, G.count, G.cap, stab, P->nbox);
}

template <int MODE, int K>
static void af_xlat(eb_afmm P, const eb_afmm_s::Group& G, const int* stab, const double2* src, double2* dst,
                    const double2* pre, const double2* post) {
  if (G.grid == 0) return;
  switch ((P->p + 3) / 4 - (K / 4 - 3)) {
    case 0: af_xlat_ks<MODE, K, K / 4 - 3>(P, G, stab, src, dst, pre, post); break;
    case 1: af_xlat_ks<MODE, K, K / 4 - 2>(P, G, stab, src, dst, pre, post); break;
    case 2: af_xlat_ks<MODE, K, K / 4 - 1>(P, G, stab, src, dst, pre, post); break;
    default: af_xlat_ks<MODE, K, K / 4>(P, G, stab, src, dst, pre, post); break;
  }
}

template <int K>
static void af_passes(eb_afmm P, double2* out, hipEvent_t* ev) {
  hipStream_t st = P->st;
  auto mark = [&](int i) {
    if (ev) (void)hipEventRecord(ev[i], st);
  };
  mark(1);
  hipLaunchKernelGGL(k_af_p2p, dim3(P->nitems), dim3(64), 0, st, P->items, P->u_off, P->u_box, P->dstart, P->dcnt,
                     P->zs, P->qa, P->qb, P->perm, P->geom, P->sgn, out);
  mark(2);
  hipLaunchKernelGGL(k_af_p2m<K>, dim3(P->nleaf), dim3(64), 0, st, P->p, P->leaves, P->dstart, P->dcnt, P->bctr,
                     P->brad, P->zs, P->qa, P->qb, P->mp);
  mark(3);
  for (const auto& g : P->m2m) af_xlat<XL_M2M, K>(P, g, P->kid, P->mp, P->mp, P->dPM, P->dQM);
  mark(4);
  af_xlat<XL_M2L, K>(P, P->m2l, P->stab27, P->mp, P->lo, P->dU, P->dNU);
  mark(5);
  if (P->nxt)
    hipLaunchKernelGGL(k_af_p2l<K>, dim3(P->nxt), dim3(64), 0, st, P->p, P->x_tgt, P->x_off, P->x_box, P->dstart,
                       P->dcnt, P->bctr, P->brad, P->zs, P->qa, P->qb, P->lo);
  mark(6);
  for (const auto& g : P->l2l) af_xlat<XL_L2L, K>(P, g, P->dpar, P->lo, P->lo, P->dPL, P->dQL);
  mark(7);
  hipLaunchKernelGGL(k_af_far<K>, dim3(P->nitems), dim3(AF_T), 0, st, P->p, P->items, P->w_off, P->w_box, P->bctr,
                     P->brad, P->zs, P->lo, P->mp, P->sgn, P->perm, P->geom, out);
  mark(8);
}

// stage timing (eb_afmm_enable_timing): events around charges, P2P, P2M, M2M, M2L, P2L, L2L, far
static std::mutex af_timing_mu;
struct AfTiming {
  hipEvent_t ev[9] = {};
  double ms[8] = {0};
};
static std::vector<std::pair<eb_afmm, AfTiming*>> af_timing;

static AfTiming* af_timer(eb_afmm P) {
  std::lock_guard<std::mutex> lk(af_timing_mu);
  for (auto& e : af_timing)
    if (e.first == P) return e.second;
  return nullptr;
}

extern "C" int eb_afmm_apply(eb_afmm P, const double* dens, double* gradw, void* stream) {
  if (!P || !dens || !gradw) return (eb_set_error("eb_afmm_apply: null argument"), EB_ERR_ARG);
  hipStream_t s = (hipStream_t)stream;
  if (s && s != P->st) {
    // order the new stream after the work queued on the old one
    hipEvent_t e;
    EB_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    EB_HIP_CHECK(hipEventRecord(e, P->st));
    EB_HIP_CHECK(hipStreamWaitEvent(s, e, 0));
    EB_HIP_CHECK(hipEventDestroy(e));
    P->st = s;
  }
  AfTiming* tm = af_timer(P);
  if (tm) {
    for (auto& e : tm->ev)
      if (!e) EB_HIP_CHECK(hipEventCreate(&e));
    EB_HIP_CHECK(hipEventRecord(tm->ev[0], P->st));
  }
  hipLaunchKernelGGL(k_charges, dim3((P->n + 255) / 256), dim3(256), 0, P->st, P->n, P->perm, (const double2*)dens,
                     P->wdzs, P->sqb, P->variant == EB_VARIANT_HILBERT ? 1 : 0, P->qa, P->qb);
  double2* out = (double2*)gradw;
  hipEvent_t* ev = tm ? tm->ev : nullptr;
  switch (P->kmax) {
    case 16: af_passes<16>(P, out, ev); break;
    case 32: af_passes<32>(P, out, ev); break;
    case 48: af_passes<48>(P, out, ev); break;
    default: af_passes<64>(P, out, ev); break;
  }
  EB_HIP_CHECK(hipGetLastError());
  if (tm) {
    EB_HIP_CHECK(hipEventSynchronize(tm->ev[8]));
    for (int i = 0; i < 8; ++i) {
      float ms = 0;
      EB_HIP_CHECK(hipEventElapsedTime(&ms, tm->ev[i], tm->ev[i + 1]));
      tm->ms[i] = ms;
    }
  }
  return EB_OK;
}

extern "C" int eb_afmm_enable_timing(eb_afmm P, int32_t on) {
  if (!P) return (eb_set_error("eb_afmm_enable_timing: null plan"), EB_ERR_ARG);
  std::lock_guard<std::mutex> lk(af_timing_mu);
  for (size_t i = 0; i < af_timing.size(); ++i)
    if (af_timing[i].first == P) {
      if (!on) {
        for (auto& e : af_timing[i].second->ev)
          if (e) (void)hipEventDestroy(e);
        delete af_timing[i].second;
        af_timing.erase(af_timing.begin() + i);
      }
      return EB_OK;
    }
  if (on) af_timing.push_back({P, new AfTiming()});
  return EB_OK;
}

extern "C" int eb_afmm_stage_ms(eb_afmm P, double* ms8) {
  AfTiming* tm = P ? af_timer(P) : nullptr;
  if (!tm || !ms8) return (eb_set_error("eb_afmm_stage_ms: timing is not enabled"), EB_ERR_ARG);
  for (int i = 0; i < 8; ++i) ms8[i] = tm->ms[i];
  return EB_OK;
}

extern "C" int eb_afmm_info(eb_afmm P, double* reals3, int64_t* counts8) {
  if (!P || !counts8) return (eb_set_error("eb_afmm_info: bad arguments"), EB_ERR_ARG);
  for (int i = 0; i < 8; ++i) counts8[i] = P->cnt[i];
  if (reals3) {
    reals3[0] = P->rscal;
    reals3[1] = P->xonew;
    reals3[2] = P->yonew;
  }
  return EB_OK;
}

extern "C" int eb_afmm_m2l_split(eb_afmm P, int64_t* out3) {
  if (!P || !out3) return (eb_set_error("eb_afmm_m2l_split: bad arguments"), EB_ERR_ARG);
  long long grp[2] = {0, 0};
  const int nb = P->nbox;
  for (int b = 0; b < nb; ++b)
    for (int o = 0; o < 27; ++o)
      if (P->stab27_h[(size_t)b * 27 + o] != nb) ++grp[o < M2L_NEAR ? 0 : 1];
  out3[0] = grp[0];
  out3[1] = grp[1];
  out3[2] = m2l_far_terms(P->p, P->kmax);
  return EB_OK;
}

// For every ordered pair (b, c) of childless boxes, the number of ways the passes carry c's
// points to b's: c in U(b); an ancestor-or-self of c in W(b); an ancestor-or-self of c in
// the interaction list of the same-level ancestor-or-self of b; c in X of an ancestor-or-self
// of b.  Every pair must count exactly once; *bad receives the pairs that do not.
extern "C" int eb_afmm_check(eb_afmm P, int64_t* bad) {
  if (!P || !bad) return (eb_set_error("eb_afmm_check: bad arguments"), EB_ERR_ARG);
  const int nb = P->nbox;
  std::vector<std::vector<int>> Us(nb), Ws(nb), Xs(nb);
  for (int b = 0; b < nb; ++b) {
    Us[b] = P->U[b];
    Ws[b] = P->W[b];
    Xs[b] = P->X[b];
    std::sort(Us[b].begin(), Us[b].end());
    std::sort(Ws[b].begin(), Ws[b].end());
    std::sort(Xs[b].begin(), Xs[b].end());
  }
  auto has = [](const std::vector<int>& v, int x) { return std::binary_search(v.begin(), v.end(), x); };
  std::vector<int> leaves;
  for (int b = 0; b < nb; ++b)
    if (P->leafv[b]) leaves.push_back(b);
  long long nbad = 0;
  std::vector<int> ab, ac;
  for (int b : leaves) {
    ab.clear();
    for (int a = b; a >= 0; a = P->par[a]) ab.push_back(a);
    for (int c : leaves) {
      ac.clear();
      for (int a = c; a >= 0; a = P->par[a]) ac.push_back(a);
      int k = has(Us[b], c) ? 1 : 0;
      for (int a : ac) k += has(Ws[b], a);
      for (int a : ab) k += has(Xs[a], c);
      for (int x : ab)
        for (int y : ac)
          if (P->lev[x] == P->lev[y] && P->lev[x] >= 2)
            for (int o = 0; o < 27; ++o) k += P->stab27_h[(size_t)x * 27 + o] == y;
      nbad += k != 1;
    }
  }
  *bad = nbad;
  return EB_OK;
}

extern "C" int eb_afmm_destroy(eb_afmm P) {
  if (P) eb_afmm_enable_timing(P, 0);
  afmm_free(P);
  return EB_OK;
}
