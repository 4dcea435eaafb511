// eb_gmres.hip -- device-resident GMRES (include/eb_gmres.h).
//
// The n-vectors (Krylov basis V, iterate, residual) live in device memory.  One Arnoldi step
// is: the operator writes A v_l into V(:, l+1); modified Gram-Schmidt runs as one fused
// kernel per basis vector (w -= h_i v_i, then the partial dot with v_{i+1}), each ending in a
// deterministic last-block reduction that leaves h_{i+1} in device memory for the next
// kernel; the Hessenberg column and the two norms then come to the host, which applies
// DORTH's re-orthogonalisation test, DHEQR's Givens rotation and ISDGMR's ITOL = 0 test in
// the reference's order, and scales v_{l+1} on the device.  DHELS and DRLCAL produce
// coefficient vectors on the host that one linear-combination kernel applies.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/eb_gmres.h"
#include "eb_internal.h"

namespace {

constexpr int GM_T = 256;
constexpr int GM_MAXBLK = 512;

struct RedBuf {
  double* part = nullptr;   // 2 * GM_MAXBLK partial sums
  unsigned* count = nullptr;
};

// Fixed-shape tree sum of one value per thread; result in thread 0.
__device__ __forceinline__ double block_tree(double v, double* s) {
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = GM_T / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  const double r = s[0];
  __syncthreads();
  return r;
}

// Writes this block's two partial sums; the last block to finish sums all partials in block
// order (the same order every run) and stores them to out0 / out1.
__device__ void finish2(double a, double b, double* __restrict__ part, unsigned* __restrict__ count,
                        double* __restrict__ out0, double* __restrict__ out1) {
  __shared__ double s[GM_T];
  __shared__ bool last;
  const double ta = block_tree(a, s), tb = block_tree(b, s);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = ta;
    part[2 * blockIdx.x + 1] = tb;
    __threadfence();
    last = atomicAdd(count, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  double pa = 0.0, pb = 0.0;
  for (int k = threadIdx.x; k < (int)gridDim.x; k += GM_T) {
    pa += __hip_atomic_load(&part[2 * k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    pb += __hip_atomic_load(&part[2 * k + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  const double ra = block_tree(pa, s), rb = block_tree(pb, s);
  if (threadIdx.x == 0) {
    if (out0) *out0 = ra;
    if (out1) *out1 = rb;
    *count = 0;
  }
}

// r = b - r (r holds A x on entry); sums |b|^2 and |r|^2
__global__ __launch_bounds__(GM_T) void k_gm_residual(int n, const double* __restrict__ b, double* __restrict__ r,
                                                      double* part, unsigned* count, double* out) {
  double sb = 0.0, sr = 0.0;
  for (int i = blockIdx.x * GM_T + threadIdx.x; i < n; i += gridDim.x * GM_T) {
    const double bi = b[i];
    const double ri = bi - r[i];
    r[i] = ri;
    sb += bi * bi;
    sr += ri * ri;
  }
  finish2(sb, sr, part, count, out, out + 1);
}

// |w|^2 and v . w
__global__ __launch_bounds__(GM_T) void k_gm_norm_dot(int n, const double* __restrict__ w,
                                                      const double* __restrict__ v, double* part, unsigned* count,
                                                      double* out_norm2, double* out_dot) {
  double s0 = 0.0, s1 = 0.0;
  for (int i = blockIdx.x * GM_T + threadIdx.x; i < n; i += gridDim.x * GM_T) {
    const double wi = w[i];
    s0 += wi * wi;
    s1 += v[i] * wi;
  }
  finish2(s0, s1, part, count, out_norm2, out_dot);
}

// v . w
__global__ __launch_bounds__(GM_T) void k_gm_dot(int n, const double* __restrict__ v, const double* __restrict__ w,
                                                 double* part, unsigned* count, double* out) {
  double s = 0.0;
  for (int i = blockIdx.x * GM_T + threadIdx.x; i < n; i += gridDim.x * GM_T) s += v[i] * w[i];
  finish2(s, 0.0, part, count, out, nullptr);
}

// DORTH's first pass (dorth.f:90-107) in one workgroup for n <= 1024 * PER: |w|^2 and h_0 =
// v_0 . w, then w -= h_{i-1} v_{i-1} and h_i = v_i . w for i = 1 .. ll (|w|^2 after the last).
// w stays in registers for the whole sweep and each basis vector is read once (v_i is kept
// for the next pass; with KEEPV false it is read again), so one launch replaces ll + 1
// launches of the multi-block kernels.  Reductions are fixed-order (wave, then block).
// Used up to MG1_NMAX unknowns (the register arrays would spill beyond 16 * 1024).
constexpr int MG1_T = 1024;
// up to this many unknowns DORTH's modified Gram-Schmidt runs in one workgroup; above it the
// classical two-pass sweep takes fewer, wider launches (about 36 us per Arnoldi step whatever
// ll, against about 3.5 us per pass in one workgroup)
constexpr int MG1_NMAX = 4 * MG1_T;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// block-wide sum over MG1_T threads, returned to every thread
__device__ __forceinline__ double block_sum1k(double v, double* s) {
  v = wave_sum(v);
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) s[wv] = v;
  __syncthreads();
  double t = lane < MG1_T / 64 ? s[lane] : 0.0;
  t = wave_sum(t);
  __syncthreads();  // s is reused by the next call
  return t;
}

template <int PER, bool KEEPV>
__global__ __launch_bounds__(MG1_T) void k_gm_mgs1(int n, int ll, double* __restrict__ V, double* __restrict__ scal,
                                                   int sh, int svn, int ssn) {
  __shared__ double s[MG1_T / 64];
  const double* v0 = V;
  double* w = V + (size_t)ll * n;
  double wr[PER], vp[KEEPV ? PER : 1];
  double a = 0.0, b = 0.0;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = threadIdx.x + k * MG1_T;
    const double wi = i < n ? w[i] : 0.0;
    const double vi = i < n ? v0[i] : 0.0;
    wr[k] = wi;
    if (KEEPV) vp[k] = vi;
    a += wi * wi;
    b += vi * wi;
  }
  a = block_sum1k(a, s);
  b = block_sum1k(b, s);
  if (threadIdx.x == 0) {
    scal[svn] = a;
    scal[sh] = b;
  }
  double h = b;
  for (int j = 1; j <= ll; ++j) {
    const double c = -h;
    const double* vprev = V + (size_t)(j - 1) * n;
    const double* vn = j < ll ? V + (size_t)j * n : nullptr;
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = threadIdx.x + k * MG1_T;
      if (i < n) {
        const double pv = KEEPV ? vp[k] : vprev[i];
        const double wi = wr[k] + c * pv;
        wr[k] = wi;
        if (vn) {
          const double vi = vn[i];
          if (KEEPV) vp[k] = vi;
          acc += vi * wi;
        } else {
          acc += wi * wi;
        }
      }
    }
    h = block_sum1k(acc, s);
    if (threadIdx.x == 0) scal[j < ll ? sh + j : ssn] = h;
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = threadIdx.x + k * MG1_T;
    if (i < n) w[i] = wr[k];
  }
}

// Above the one-workgroup limit, one Arnoldi step orthogonalises by classical Gram-Schmidt
// with one re-orthogonalisation: two passes of (the ll dots in chunks; their fixed-order sums
// and the update), four launches per step instead of ll + 1.  The first pass's dots also
// give |w|^2 before, the second pass's update |w|^2 after.
constexpr int CGS_CH = 2048;

__global__ __launch_bounds__(GM_T) void k_gm_cgs_dot(int n, int nch, int ll, const double* __restrict__ V,
                                                     const double* __restrict__ w, double* __restrict__ part) {
  __shared__ double s[GM_T];
  const int c = blockIdx.x, j = blockIdx.y;
  // grid row j < ll: v_j . w; row ll (launched on the first pass only): |w|^2
  const double* v = j < ll ? V + (size_t)j * n : w;
  const int i0 = c * CGS_CH, i1 = min(n, i0 + CGS_CH);
  double a = 0.0;
  for (int i = i0 + (int)threadIdx.x; i < i1; i += GM_T) a += v[i] * w[i];
  a = block_tree(a, s);
  if (threadIdx.x == 0) part[(size_t)j * nch + c] = a;
}

// One classical pass's update: every block first forms h_j = sum over chunks of part (in
// chunk order, so every block holds the same h) in LDS, then w -= sum_j h_j V(:, j), j in
// order.  Block 0 stores scal[sh + j] = h_j on the first pass (+= on the second) and, when vn
// is given, |w|^2 before the pass from the partials' row ll.  With NORM the blocks then sum
// |w|^2 after the update (finish2, fixed order) into *sn.  Dynamic LDS: ll doubles.
template <bool NORM>
__global__ __launch_bounds__(GM_T) void k_gm_cgs_update(int n, int ll, int nch, const double* __restrict__ V,
                                                        const double* __restrict__ part, double* __restrict__ scal,
                                                        int sh, int second, double* __restrict__ vn,
                                                        double* __restrict__ w, double* red_part, unsigned* red_count,
                                                        double* __restrict__ sn) {
  extern __shared__ double hs[];
  // h_j = the chunk partials summed in chunk order (one thread per row), eight loads in flight
  // at a time: summed one by one, each chunk waited a memory latency
  const int nj = vn ? ll + 1 : ll;
  for (int j = threadIdx.x; j < nj; j += GM_T) {
    const double* pj = part + (size_t)j * nch;
    double h = 0.0;
    int c = 0;
    for (; c + 8 <= nch; c += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = pj[c + u];
#pragma unroll
      for (int u = 0; u < 8; ++u) h += v[u];
    }
    for (; c < nch; ++c) h += pj[c];
    if (j < ll) {
      hs[j] = h;
      if (blockIdx.x == 0) scal[sh + j] = second ? scal[sh + j] + h : h;
    } else if (blockIdx.x == 0) {
      *vn = h;
    }
  }
  __syncthreads();
  double s2 = 0.0;
  for (int i = blockIdx.x * GM_T + threadIdx.x; i < n; i += gridDim.x * GM_T) {
    double t = w[i];
    for (int j = 0; j < ll; ++j) t -= hs[j] * V[(size_t)j * n + i];
    w[i] = t;
    s2 += t * t;
  }
  if (NORM) finish2(s2, 0.0, red_part, red_count, sn, nullptr);
}

// w += c v (c from the host)
__global__ void k_gm_axpy(int n, double c, const double* __restrict__ v, double* __restrict__ w) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) w[i] += c * v[i];
}

__global__ void k_gm_scale(int n, double c, double* __restrict__ v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] *= c;
}

__global__ void k_gm_copy(int n, const double* __restrict__ a, double* __restrict__ b) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) b[i] = a[i];
}

// z = sum_j c_j V(:, j), j = 0..m-1 in order (the DAXPY sequence of dpigmr.f:411-413);
// then x += z when x is non-NULL, else out = z
__global__ void k_gm_combine(int n, int m, const double* __restrict__ V, size_t ldv, const double* __restrict__ c,
                             double* __restrict__ x, double* __restrict__ out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double z = 0.0;
    for (int j = 0; j < m; ++j) z += c[j] * V[(size_t)j * ldv + i];
    if (x) x[i] += z;
    else out[i] = z;
  }
}

// y = A x, A column-major n x n
__global__ void k_gm_dense(int n, const double* __restrict__ a, const double* __restrict__ x, double* __restrict__ y) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < n; ++j) s += a[(size_t)j * n + i] * x[j];
    y[i] = s;
  }
}

}  // namespace

struct eb_gmres_s {
  int n = 0, maxl = 0;
  hipStream_t st = nullptr;
  double* V = nullptr;        // n x (maxl + 1), column-major
  double* r = nullptr;        // residual / scratch (n)
  double* scal = nullptr;     // device scalars: h[0..maxl], norms
  double* coef = nullptr;     // combination coefficients (maxl + 1)
  RedBuf red;
  int nblk = 1;
  int nch = 0;                // chunks of the classical Gram-Schmidt passes (n > MG1 limit)
  double* cgs_part = nullptr; // (maxl + 1) x nch partial dots (the last row: |w|^2)
  double* hpin = nullptr;     // pinned host copy of scal (the per-iteration read-back)
};

namespace {

constexpr int S_H = 0;  // scal[S_H + i] = h_i (i = 0..maxl-1), then the slots below
inline int s_vn(const eb_gmres_s* g) { return g->maxl + 1; }
inline int s_sn(const eb_gmres_s* g) { return g->maxl + 2; }
inline int s_aux(const eb_gmres_s* g) { return g->maxl + 3; }  // 2 slots

int grid_for(int n) { return std::max(1, std::min(GM_MAXBLK, (n + GM_T * 4 - 1) / (GM_T * 4))); }
int ew_grid(int n) { return std::max(1, std::min(4096, (n + 255) / 256)); }

// DHEQR column update for column ll (1-based), dheqr.f:100-176 (IJOB > 1, or the full
// factorisation of a 2x1 matrix for ll = 1 which is the same arithmetic on one column).
// hes is column-major with leading dimension ldh; q holds (c, s) pairs.  Returns INFO.
__host__ __device__ int dheqr_column(double* hes, int ldh, int ll, double* q) {
  auto A = [&](int i, int j) -> double& { return hes[(size_t)(j - 1) * ldh + (i - 1)]; };
  for (int k = 1; k <= ll - 1; ++k) {
    const int i = 2 * (k - 1);
    const double t1 = A(k, ll), t2 = A(k + 1, ll), c = q[i], s = q[i + 1];
    A(k, ll) = c * t1 - s * t2;
    A(k + 1, ll) = s * t1 + c * t2;
  }
  const double t1 = A(ll, ll), t2 = A(ll + 1, ll);
  double c, s;
  if (t2 == 0.0) {
    c = 1.0;
    s = 0.0;
  } else if (fabs(t2) >= fabs(t1)) {
    const double t = t1 / t2;
    s = -1.0 / sqrt(1.0 + t * t);
    c = -s * t;
  } else {
    const double t = t2 / t1;
    c = 1.0 / sqrt(1.0 + t * t);
    s = -c * t;
  }
  q[2 * (ll - 1)] = c;
  q[2 * (ll - 1) + 1] = s;
  A(ll, ll) = c * t1 - s * t2;
  return A(ll, ll) == 0.0 ? ll : 0;
}

// DHELS (dhels.f): least-squares solution of the factored Hessenberg system, in place on b.
__host__ __device__ void dhels(const double* hes, int ldh, int n, const double* q, double* b) {
  auto A = [&](int i, int j) { return hes[(size_t)(j - 1) * ldh + (i - 1)]; };
  for (int k = 1; k <= n; ++k) {
    const double c = q[2 * (k - 1)], s = q[2 * (k - 1) + 1];
    const double t1 = b[k - 1], t2 = b[k];
    b[k - 1] = c * t1 - s * t2;
    b[k] = s * t1 + c * t2;
  }
  for (int kb = 1; kb <= n; ++kb) {
    const int k = n + 1 - kb;
    b[k - 1] = b[k - 1] / A(k, k);
    const double t = -b[k - 1];
    for (int i = 1; i <= k - 1; ++i) b[i - 1] = b[i - 1] + t * A(i, k);
  }
}

// The scalars go through a pinned buffer: a copy into pageable memory is staged by the
// runtime and costs several microseconds more per Arnoldi step.
int fetch(eb_gmres_s* g, int first, int count, double* host) {
  EB_HIP_CHECK(hipMemcpyAsync(g->hpin, g->scal + first, sizeof(double) * count, hipMemcpyDeviceToHost, g->st));
  EB_HIP_CHECK(hipStreamSynchronize(g->st));
  memcpy(host, g->hpin, sizeof(double) * count);
  return EB_OK;
}

double* col(eb_gmres_s* g, int j) { return g->V + (size_t)j * g->n; }  // 0-based column

// One DPIGMR cycle (dpigmr.f:295-437) from the residual already in r.  Outputs as DPIGMR:
// lgmr, iflag, rhol, err; adds the correction to x; when a restart follows, r receives the
// new residual (DRLCAL).
int dpigmr_cycle(eb_gmres_s* g, eb_apply_fn apply, void* ctx, double* x, double r0nrm, double bnrm, double tol,
                 int nrsts, int nrmax, int& lgmr, int& iflag, double& rhol, double& err) {
  const int n = g->n, maxl = g->maxl, ldh = maxl + 1;
  hipStream_t st = g->st;
  iflag = 0;
  lgmr = 0;
  rhol = r0nrm;
  // ISDGMR at the start of the cycle (ITOL = 0): ERR = RNRM / BNRM
  err = r0nrm / bnrm;
  if (err <= tol) return EB_OK;  // converged with a zero correction
  hipLaunchKernelGGL(k_gm_copy, dim3(ew_grid(n)), dim3(256), 0, st, n, g->r, col(g, 0));
  hipLaunchKernelGGL(k_gm_scale, dim3(ew_grid(n)), dim3(256), 0, st, n, 1.0 / r0nrm, col(g, 0));
  std::vector<double> hes((size_t)ldh * maxl, 0.0), q(2 * maxl, 0.0), h(maxl + 3);
  double prod = 1.0, rho = r0nrm, snormw = 0.0;
  const int nb = g->nblk;
  bool converged = false;
  int ll = 0;
  for (ll = 1; ll <= maxl; ++ll) {
    lgmr = ll;
    double* w = col(g, ll);
    int rc = apply(ctx, col(g, ll - 1), w, st);
    if (rc != EB_OK) return rc;
    // DORTH (dorth.f:90-121): VNRM, then modified Gram-Schmidt against v_1 .. v_ll
    if (n <= MG1_NMAX) {
      hipLaunchKernelGGL((k_gm_mgs1<MG1_NMAX / MG1_T, true>), dim3(1), dim3(MG1_T), 0, st, n, ll, g->V, g->scal, S_H,
                         s_vn(g), s_sn(g));
    } else {
      // two classical passes against v_0 .. v_{ll-1} in four launches: the first pass's dots
      // carry |w|^2 before (grid row ll), the second pass's update |w|^2 after
      const size_t lds = sizeof(double) * (size_t)(ll + 1);
      hipLaunchKernelGGL(k_gm_cgs_dot, dim3(g->nch, ll + 1), dim3(GM_T), 0, st, n, g->nch, ll, g->V, w,
                         g->cgs_part);
      // the first pass's update has no norm to finish, so it takes one element per thread
      const int nbu = std::max(1, std::min(GM_MAXBLK, (n + GM_T - 1) / GM_T));
      hipLaunchKernelGGL(k_gm_cgs_update<false>, dim3(nbu), dim3(GM_T), lds, st, n, ll, g->nch, g->V, g->cgs_part,
                         g->scal, S_H, 0, g->scal + s_vn(g), w, nullptr, nullptr, nullptr);
      hipLaunchKernelGGL(k_gm_cgs_dot, dim3(g->nch, ll), dim3(GM_T), 0, st, n, g->nch, ll, g->V, w, g->cgs_part);
      hipLaunchKernelGGL(k_gm_cgs_update<true>, dim3(nb), dim3(GM_T), lds, st, n, ll, g->nch, g->V, g->cgs_part,
                         g->scal, S_H, 1, nullptr, w, g->red.part, g->red.count, g->scal + s_sn(g));
    }
    EB_HIP_CHECK(hipGetLastError());
    if ((rc = fetch(g, 0, maxl + 3, h.data())) != EB_OK) return rc;
    const double vnrm = sqrt(h[s_vn(g)]);
    snormw = sqrt(h[s_sn(g)]);
    for (int i = 1; i <= ll; ++i) hes[(size_t)(ll - 1) * ldh + (i - 1)] = h[S_H + i - 1];
    if (!(vnrm + 0.001 * snormw != vnrm)) {
      // re-orthogonalisation pass (dorth.f:108-121)
      double sumdsq = 0.0;
      for (int i = 1; i <= ll; ++i) {
        hipLaunchKernelGGL(k_gm_dot, dim3(nb), dim3(GM_T), 0, st, n, col(g, i - 1), w, g->red.part, g->red.count,
                           g->scal + s_aux(g));
        double d;
        if ((rc = fetch(g, s_aux(g), 1, &d)) != EB_OK) return rc;
        const double tem = -d;
        double& hil = hes[(size_t)(ll - 1) * ldh + (i - 1)];
        if (hil + 0.001 * tem == hil) continue;
        hil = hil - tem;
        hipLaunchKernelGGL(k_gm_axpy, dim3(ew_grid(n)), dim3(256), 0, st, n, tem, col(g, i - 1), w);
        sumdsq = sumdsq + tem * tem;
      }
      if (sumdsq != 0.0) {
        const double arg = std::max(0.0, snormw * snormw - sumdsq);
        snormw = sqrt(arg);
      }
    }
    hes[(size_t)(ll - 1) * ldh + ll] = snormw;  // HES(LL+1, LL)
    const int info = dheqr_column(hes.data(), ldh, ll, q.data());
    if (info == ll) {  // singular Hessenberg: no progress (dpigmr.f:366, 419-423)
      iflag = 2;
      return EB_OK;
    }
    prod = prod * q[2 * ll - 1];
    rho = fabs(prod * r0nrm);
    rhol = rho;
    err = rho / bnrm;  // ISDGMR, ITOL = 0
    if (err <= tol) {
      converged = true;
      break;
    }
    if (ll == maxl) break;
    hipLaunchKernelGGL(k_gm_scale, dim3(ew_grid(n)), dim3(256), 0, st, n, 1.0 / snormw, w);
  }
  if (!converged) {
    if (!(rho < r0nrm)) {
      iflag = 2;
      return EB_OK;
    }
    iflag = 1;
  }
  // DHELS on (R0NRM, 0, ..., 0), then z = sum r0(i) v_i and x += z (dpigmr.f:400-415)
  const int l = lgmr;
  std::vector<double> b(l + 1, 0.0);
  b[0] = r0nrm;
  dhels(hes.data(), ldh, l, q.data(), b.data());
  if (iflag == 1 && nrmax > 0) {
    // DRLCAL (drlcal.f, KMP = MAXL): the residual of the new iterate as a combination of
    // v_1 .. v_{l+1}, for the next cycle; v_{l+1} is unscaled, hence C / SNORMW.
    std::vector<double> a(l + 1, 0.0);
    a[0] = 1.0;
    for (int i = 1; i <= l - 1; ++i) {
      const double s = q[2 * i - 1], c = q[2 * i - 2];
      for (int j = 0; j < i; ++j) a[j] *= s;
      a[i] = c;
    }
    const double s = q[2 * l - 1], c = q[2 * l - 2] / snormw;
    for (int j = 0; j < l; ++j) a[j] *= s;
    a[l] = c;
    const double tem = r0nrm * prod;
    for (double& v : a) v *= tem;
    EB_HIP_CHECK(hipMemcpyAsync(g->coef, a.data(), sizeof(double) * (l + 1), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_gm_combine, dim3(ew_grid(n)), dim3(256), 0, st, n, l + 1, g->V, (size_t)n, g->coef,
                       nullptr, g->r);
    EB_HIP_CHECK(hipStreamSynchronize(st));  // coef is rewritten below
  }
  EB_HIP_CHECK(hipMemcpyAsync(g->coef, b.data(), sizeof(double) * l, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_gm_combine, dim3(ew_grid(n)), dim3(256), 0, st, n, l, g->V, (size_t)n, g->coef, x, nullptr);
  EB_HIP_CHECK(hipGetLastError());
  EB_HIP_CHECK(hipStreamSynchronize(st));
  (void)nrsts;
  return EB_OK;
}

}  // namespace

extern "C" int eb_gmres_create(eb_gmres* out, int32_t n, int32_t maxl, void* stream) {
  if (!out || n < 1) return (eb_set_error("eb_gmres_create: bad arguments"), EB_ERR_ARG);
  *out = nullptr;
  if (maxl == 0) maxl = 10;  // dgmres.f: MAXL = IGWK(1), 0 -> 10
  if (maxl > n) maxl = n;
  if (maxl < 1) return (eb_set_error("eb_gmres_create: MAXL = %d", maxl), EB_ERR_ARG);
  eb_gmres g = new eb_gmres_s();
  g->n = n;
  g->maxl = maxl;
  g->st = (hipStream_t)stream;
  g->nblk = grid_for(n);
  bool ok = hipMalloc((void**)&g->V, sizeof(double) * (size_t)n * (maxl + 1)) == hipSuccess &&
            hipMalloc((void**)&g->r, sizeof(double) * (size_t)n) == hipSuccess &&
            hipMalloc((void**)&g->scal, sizeof(double) * (maxl + 8)) == hipSuccess &&
            hipMalloc((void**)&g->coef, sizeof(double) * (maxl + 2)) == hipSuccess &&
            hipMalloc((void**)&g->red.part, sizeof(double) * 2 * GM_MAXBLK) == hipSuccess &&
            hipMalloc((void**)&g->red.count, sizeof(unsigned)) == hipSuccess &&
            hipHostMalloc((void**)&g->hpin, sizeof(double) * (maxl + 8), hipHostMallocDefault) == hipSuccess;
  if (ok && n > MG1_NMAX) {
    g->nch = (n + CGS_CH - 1) / CGS_CH;
    ok = hipMalloc((void**)&g->cgs_part, sizeof(double) * (size_t)(maxl + 1) * g->nch) == hipSuccess;
  }
  if (!ok || hipMemsetAsync(g->red.count, 0, sizeof(unsigned), g->st) != hipSuccess ||
      hipMemsetAsync(g->scal, 0, sizeof(double) * (maxl + 8), g->st) != hipSuccess) {
    eb_gmres_destroy(g);
    return (eb_set_error("eb_gmres_create: device allocation failed"), EB_ERR_NOMEM);
  }
  *out = g;
  return EB_OK;
}

extern "C" int eb_gmres_destroy(eb_gmres g) {
  if (!g) return EB_OK;
  if (g->st) (void)hipStreamSynchronize(g->st);
  void* p[] = {g->V, g->r, g->scal, g->coef, g->red.part, g->red.count, g->cgs_part};
  for (void* q : p)
    if (q) (void)hipFree(q);
  if (g->hpin) (void)hipHostFree(g->hpin);
  delete g;
  return EB_OK;
}

extern "C" int eb_gmres_solve(eb_gmres g, eb_apply_fn apply, void* ctx, const double* b, double* x, double tol,
                              int32_t nrmax, int32_t* iter, double* err, int32_t* ierr) {
  if (!g || !apply || !b || !x) return (eb_set_error("eb_gmres_solve: bad arguments"), EB_ERR_ARG);
  const int n = g->n;
  hipStream_t st = g->st;
  if (nrmax == 0) nrmax = 10;  // dgmres.f: IGWK(5) 0 -> 10, -1 -> 0
  if (nrmax == -1) nrmax = 0;
  if (tol == 0.0) tol = 500.0 * 1.1102230246251565e-16;  // 500 * D1MACH(3) = 500 * 2^-53
  int it = 0, ier = 0;
  double e = 0.0, rhol = 0.0;
  // r = b - A x; BNRM = |b|
  int rc = apply(ctx, x, g->r, st);
  if (rc != EB_OK) return rc;
  hipLaunchKernelGGL(k_gm_residual, dim3(g->nblk), dim3(GM_T), 0, st, n, b, g->r, g->red.part, g->red.count,
                     g->scal + s_aux(g));
  EB_HIP_CHECK(hipGetLastError());
  double nrm2[2];
  if ((rc = fetch(g, s_aux(g), 2, nrm2)) != EB_OK) return rc;
  const double bnrm = sqrt(nrm2[0]);
  double r0nrm = sqrt(nrm2[1]);
  for (int nrsts = 0;; ++nrsts) {
    if (nrsts > nrmax) {
      ier = 1;
      break;
    }
    int lgmr = 0, iflag = 0;
    rc = dpigmr_cycle(g, apply, ctx, x, r0nrm, bnrm, tol, nrsts, nrmax, lgmr, iflag, rhol, e);
    if (rc != EB_OK) return rc;
    it += lgmr;
    if (iflag == 0) {
      ier = 0;
      break;
    }
    if (iflag == 2) {
      ier = 2;
      break;
    }
    // restart from the DRLCAL residual now in r
    if (nrsts + 1 > nrmax) {
      ier = 1;
      break;
    }
    hipLaunchKernelGGL(k_gm_norm_dot, dim3(g->nblk), dim3(GM_T), 0, st, n, g->r, g->r, g->red.part, g->red.count,
                       g->scal + s_aux(g), nullptr);
    double nr;
    if ((rc = fetch(g, s_aux(g), 1, &nr)) != EB_OK) return rc;
    r0nrm = sqrt(nr);
  }
  if (iter) *iter = it;
  if (err) *err = e;
  if (ierr) *ierr = ier;
  return EB_OK;
}

extern "C" int eb_dense_gmres_apply(void* op, const double* x, double* y, void* stream) {
  const eb_dense_op* d = (const eb_dense_op*)op;
  if (!d || !d->a || d->n < 1) return (eb_set_error("eb_dense_gmres_apply: bad operator"), EB_ERR_ARG);
  hipLaunchKernelGGL(k_gm_dense, dim3(std::max(1, std::min(1024, (d->n + 255) / 256))), dim3(256), 0,
                     (hipStream_t)stream, d->n, d->a, x, y);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" int eb_matvec_gmres_apply(void* mv, const double* x, double* y, void* stream) {
  return eb_matvec_apply((eb_matvec)mv, x, y, nullptr, nullptr, nullptr, stream);
}

extern "C" int eb_stokes_solve(eb_matvec mv, eb_gmres g, const double* rhs, double* soln, double* za, double* pk,
                               int32_t* iter, double* err, int32_t* ierr, void* stream) {
  if (!mv || !g || !rhs || !soln) return (eb_set_error("eb_stokes_solve: bad arguments"), EB_ERR_ARG);
  hipStream_t st = (hipStream_t)stream;
  if (st != g->st) {
    eb_set_error("eb_stokes_solve: the GMRES workspace was created on another stream");
    return EB_ERR_ARG;
  }
  // initial guess soln = rhs (stokesbubb.f:343-346), ITOL 0, TOL 1e-10, IGWK(5) = -1
  hipLaunchKernelGGL(k_gm_copy, dim3(ew_grid(g->n)), dim3(256), 0, st, g->n, rhs, soln);
  int rc = eb_gmres_solve(g, eb_matvec_gmres_apply, mv, rhs, soln, 1.0e-10, -1, iter, err, ierr);
  if (rc != EB_OK) return rc;
  if (za || pk) rc = eb_matvec_parts(mv, soln, za, pk, stream);
  return rc;
}

// ==========================================================================================
// Batched GMRES: many independent systems A_s x_s = b_s in lockstep
// ==========================================================================================
//
// System s owns entries [off[s], off[s+1]) of every n-vector; the operator applies all
// systems at once (a block-diagonal A).  Each system follows DGMRES with IGWK(5) = -1 (one
// DPIGMR cycle, no restart), SOLVEI's setting (stokesbubb.f:294-415): its own Hessenberg
// matrix, rotations, norms and tests, in the reference's order, on the host.  On the device
// one workgroup owns one system, so every reduction is a fixed tree inside one block and one
// kernel runs the whole modified Gram-Schmidt step of all systems.  Systems that have
// converged or stopped drop out of the kernels; their Krylov columns stay untouched until
// the final linear combination, which updates all iterates in one launch.
namespace {

constexpr int GB_T = 256;

__device__ __forceinline__ double2 block_sum2(double a, double b, double2* s) {
  s[threadIdx.x] = make_double2(a, b);
  __syncthreads();
  for (int o = GB_T / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s[threadIdx.x].x += s[threadIdx.x + o].x;
      s[threadIdx.x].y += s[threadIdx.x + o].y;
    }
    __syncthreads();
  }
  const double2 r = s[0];
  __syncthreads();
  return r;
}

// per-system slots of the scalar array (hld = maxl + 4 doubles per system)
// [0, maxl): Hessenberg column; maxl: |w|^2 before MGS; maxl+1: |w|^2 after; maxl+2: SUMDSQ;
// maxl+3: spare.  The residual pass uses slots 0 and 1 for |b|^2 and |r|^2.

// r = b - r (r holds A x); |b|^2 and |r|^2 per system
__global__ __launch_bounds__(GB_T) void k_gmb_residual(const int* __restrict__ off, const double* __restrict__ b,
                                                       double* __restrict__ r, double* __restrict__ hs, int hld) {
  __shared__ double2 red[GB_T];
  const int s = blockIdx.x, o = off[s], n = off[s + 1] - o;
  double sb = 0.0, sr = 0.0;
  for (int i = threadIdx.x; i < n; i += GB_T) {
    const double bi = b[o + i];
    const double ri = bi - r[o + i];
    r[o + i] = ri;
    sb += bi * bi;
    sr += ri * ri;
  }
  const double2 t = block_sum2(sb, sr, red);
  if (threadIdx.x == 0) {
    hs[(size_t)s * hld] = t.x;
    hs[(size_t)s * hld + 1] = t.y;
  }
}

// dst = src * fac[s] for the systems with act[s]
__global__ void k_gmb_scale(const int* __restrict__ off, const int* __restrict__ act, const double* __restrict__ fac,
                            const double* __restrict__ src, double* __restrict__ dst) {
  const int s = blockIdx.x;
  if (!act[s]) return;
  const int o = off[s], n = off[s + 1] - o;
  const double c = fac[s];
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[o + i] = src[o + i] * c;
}

// DORTH's first pass (dorth.f:90-107) on w = V(:, ll) of every active system: |w|^2, then
// h_i = v_i . w and w -= h_i v_i for i = 0 .. ll-1 (MGS), then |w|^2
__device__ __forceinline__ void gmb_orth_body(int s, int ll, const int* __restrict__ off, double* __restrict__ V,
                                              size_t ldv, double* __restrict__ hs, int hld, int maxl, double2* red) {
  const int o = off[s], n = off[s + 1] - o;
  double* w = V + (size_t)ll * ldv + o;
  double* h = hs + (size_t)s * hld;
  const double* v0 = V + o;
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < n; i += GB_T) {
    const double wi = w[i];
    a += wi * wi;
    b += v0[i] * wi;
  }
  double2 t = block_sum2(a, b, red);
  if (threadIdx.x == 0) {
    h[maxl] = t.x;
    h[0] = t.y;
  }
  double hprev = t.y;
  for (int i = 1; i <= ll; ++i) {
    const double c = -hprev;
    const double* vi = V + (size_t)(i - 1) * ldv + o;
    const double* vn = V + (size_t)i * ldv + o;
    double acc = 0.0;
    if (i < ll) {
      for (int j = threadIdx.x; j < n; j += GB_T) {
        const double wj = w[j] + c * vi[j];
        w[j] = wj;
        acc += vn[j] * wj;
      }
    } else {
      for (int j = threadIdx.x; j < n; j += GB_T) {
        const double wj = w[j] + c * vi[j];
        w[j] = wj;
        acc += wj * wj;
      }
    }
    t = block_sum2(acc, 0.0, red);
    hprev = t.x;
    if (threadIdx.x == 0) h[i < ll ? i : maxl + 1] = t.x;
  }
}

__global__ __launch_bounds__(GB_T) void k_gmb_orth(int ll, const int* __restrict__ off, const int* __restrict__ act,
                                                   double* __restrict__ V, size_t ldv, double* __restrict__ hs,
                                                   int hld, int maxl) {
  __shared__ double2 red[GB_T];
  const int s = blockIdx.x;
  if (!act[s]) return;
  gmb_orth_body(s, ll, off, V, ldv, hs, hld, maxl, red);
}

// DORTH's re-orthogonalisation pass (dorth.f:108-121) for the flagged systems: the updated
// Hessenberg column stays in h[0..ll) and SUMDSQ goes to h[maxl + 2]
__device__ __forceinline__ void gmb_reorth_body(int s, int ll, const int* __restrict__ off, double* __restrict__ V,
                                                size_t ldv, double* __restrict__ hs, int hld, int maxl, double2* red) {
  const int o = off[s], n = off[s + 1] - o;
  double* w = V + (size_t)ll * ldv + o;
  double* h = hs + (size_t)s * hld;
  double sumdsq = 0.0;
  for (int i = 1; i <= ll; ++i) {
    const double* vi = V + (size_t)(i - 1) * ldv + o;
    double acc = 0.0;
    for (int j = threadIdx.x; j < n; j += GB_T) acc += vi[j] * w[j];
    const double tem = -block_sum2(acc, 0.0, red).x;
    const double hil = h[i - 1];
    if (hil + 0.001 * tem == hil) continue;
    for (int j = threadIdx.x; j < n; j += GB_T) w[j] += tem * vi[j];
    __syncthreads();  // every lane has read h[i - 1] before it changes
    if (threadIdx.x == 0) h[i - 1] = hil - tem;
    sumdsq = sumdsq + tem * tem;
  }
  if (threadIdx.x == 0) h[maxl + 2] = sumdsq;
}

__global__ __launch_bounds__(GB_T) void k_gmb_reorth(int ll, const int* __restrict__ off, const int* __restrict__ flag,
                                                     double* __restrict__ V, size_t ldv, double* __restrict__ hs,
                                                     int hld, int maxl) {
  __shared__ double2 red[GB_T];
  const int s = blockIdx.x;
  if (!flag[s]) return;
  gmb_reorth_body(s, ll, off, V, ldv, hs, hld, maxl, red);
}

// ---- device-side control (no host synchronisation inside a solve) ---------------------------
// The scalar work of one Arnoldi step of every system -- DORTH's re-orthogonalisation test,
// DHEQR's rotation, ISDGMR's test and at the end DHELS -- as the host loop below does it, in
// the same order and arithmetic, one thread per system.  With it a solve is a fixed sequence
// of launches that a hipGraph can hold.

struct GmbState {
  double bnrm, r0nrm, prod, rho, err, snormw;
  int it, ie;
};

__global__ void k_gmb_dstart(int S, double tol, const double* __restrict__ hs, int hld, int* __restrict__ act,
                             double* __restrict__ fac, int* __restrict__ lcnt, GmbState* __restrict__ st) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const double* h = hs + (size_t)s * hld;
  GmbState g;
  g.bnrm = sqrt(h[0]);
  g.r0nrm = sqrt(h[1]);
  g.err = g.r0nrm / g.bnrm;  // ISDGMR, ITOL = 0, at the start of the cycle
  g.prod = 1.0;
  g.rho = 0.0;
  g.snormw = 0.0;
  g.it = 0;
  g.ie = 0;
  const int a = g.err <= tol ? 0 : 1;
  act[s] = a;
  fac[s] = a ? 1.0 / g.r0nrm : 0.0;
  lcnt[s] = 0;
  st[s] = g;
}

// after the orthogonalisation pass: |w| after MGS, and DORTH's test for a second pass
// (dorth.f:104-107)
// one active system: returns the re-orthogonalisation flag
__device__ __forceinline__ int gmb_dctl1_one(int s, const double* __restrict__ hs, int hld, int mx,
                                            GmbState* __restrict__ st) {
  const double* h = hs + (size_t)s * hld;
  const double vnrm = sqrt(h[mx]);
  const double snormw = sqrt(h[mx + 1]);
  st[s].snormw = snormw;
  return !(vnrm + 0.001 * snormw != vnrm) ? 1 : 0;
}



// after the re-orthogonalisation pass: the Hessenberg column, DHEQR, the tests of dpigmr.f:366-392 and, for a
// system that stops with an update, DHELS into cf (stride cld).  A system still iterating
// after `budget` steps stops with the update of those steps and IE = 3 (counted in *fail).
// one active system (act[s] = 1 on entry); fac[s] is 0 on entry
__device__ void gmb_dctl2_one(int s, int ll, double tol, int budget, const double* __restrict__ hs, int hld, int mx,
                              const int* __restrict__ maxl_s, int* __restrict__ act, const int* __restrict__ flag,
                              double* __restrict__ fac, int* __restrict__ lcnt, double* __restrict__ cf, int cld,
                              double* __restrict__ hesa, double* __restrict__ qa, GmbState* __restrict__ st,
                              int* __restrict__ fail) {
  const double* h = hs + (size_t)s * hld;
  const int ldh = mx + 1;
  GmbState g = st[s];
  if (flag[s]) {
    const double sumdsq = h[mx + 2];
    if (sumdsq != 0.0) g.snormw = sqrt(fmax(0.0, g.snormw * g.snormw - sumdsq));
  }
  double* hes = hesa + (size_t)s * ldh * mx;
  double* q = qa + (size_t)s * 2 * mx;
  for (int i = 1; i <= ll; ++i) hes[(size_t)(ll - 1) * ldh + (i - 1)] = h[i - 1];
  hes[(size_t)(ll - 1) * ldh + ll] = g.snormw;
  const int info = dheqr_column(hes, ldh, ll, q);
  g.it = ll;
  int done = 0;
  if (info == ll) {
    done = 2;
    g.ie = 2;
  } else {
    g.prod = g.prod * q[2 * ll - 1];
    g.rho = fabs(g.prod * g.r0nrm);
    g.err = g.rho / g.bnrm;
    if (g.err <= tol) {
      done = 1;
      g.ie = 0;
    } else if (ll == maxl_s[s]) {
      if (!(g.rho < g.r0nrm)) {
        done = 2;
        g.ie = 2;
      } else {
        done = 1;
        g.ie = 1;
      }
    } else if (ll == budget) {
      done = 1;
      g.ie = 3;
      atomicAdd(fail, 1);
    }
  }
  if (done) {
    act[s] = 0;
    if (done == 1) {
      double* bb = cf + (size_t)s * cld;
      bb[0] = g.r0nrm;
      for (int j = 1; j <= ll; ++j) bb[j] = 0.0;
      dhels(hes, ldh, ll, q, bb);
      lcnt[s] = ll;
    }
  } else {
    fac[s] = 1.0 / g.snormw;
  }
  st[s] = g;
}



// One Arnoldi step of every system after its operator call, in one launch: one workgroup per
// system runs the orthogonalisation pass (gmb_orth_body), DORTH's test (gmb_dctl1_one, thread
// 0), the re-orthogonalisation pass if flagged, the rotation and stopping tests (gmb_dctl2_one,
// thread 0) and the scaling of the new basis vector, in that order and arithmetic.
// Systems are independent, so nothing orders one system's step after another's.
__global__ __launch_bounds__(GB_T) void k_gmb_step(int ll, double tol, int budget, const int* __restrict__ off,
                                                   double* __restrict__ V, size_t ldv, double* __restrict__ hs,
                                                   int hld, int mx, const int* __restrict__ maxl_s,
                                                   int* __restrict__ act, int* __restrict__ flag,
                                                   double* __restrict__ fac, int* __restrict__ lcnt,
                                                   double* __restrict__ cf, int cld, double* __restrict__ hesa,
                                                   double* __restrict__ qa, GmbState* __restrict__ st,
                                                   int* __restrict__ fail) {
  __shared__ double2 red[GB_T];
  __shared__ int sh_flag, sh_act;
  __shared__ double sh_fac;
  const int s = blockIdx.x;
  if (!act[s]) {  // the whole block: before any barrier
    if (threadIdx.x == 0) {
      flag[s] = 0;
      fac[s] = 0.0;
    }
    return;
  }
  gmb_orth_body(s, ll, off, V, ldv, hs, hld, mx, red);
  if (threadIdx.x == 0) {
    const int f = gmb_dctl1_one(s, hs, hld, mx, st);
    flag[s] = f;
    sh_flag = f;
  }
  __syncthreads();
  if (sh_flag) gmb_reorth_body(s, ll, off, V, ldv, hs, hld, mx, red);
  __syncthreads();
  if (threadIdx.x == 0) {
    fac[s] = 0.0;
    gmb_dctl2_one(s, ll, tol, budget, hs, hld, mx, maxl_s, act, flag, fac, lcnt, cf, cld, hesa, qa, st, fail);
    sh_act = act[s];
    sh_fac = fac[s];
  }
  __syncthreads();
  if (sh_act) {
    const int o = off[s], n = off[s + 1] - o;
    double* w = V + (size_t)ll * ldv + o;
    const double c = sh_fac;
    for (int i = threadIdx.x; i < n; i += GB_T) w[i] *= c;
  }
}

__global__ void k_gmb_zero(int* __restrict__ v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *v = 0;
}

// every system active: the operator must form the initial residual of all of them
__global__ void k_gmb_all_active(int S, int* __restrict__ act) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < S) act[s] = 1;
}

// x_s += sum_{j < l_s} c_s[j] V_s(:, j) in the order of dpigmr.f:411-413
__global__ void k_gmb_combine(const int* __restrict__ off, const int* __restrict__ lcnt, const double* __restrict__ coef,
                              int cld, const double* __restrict__ V, size_t ldv, double* __restrict__ x) {
  const int s = blockIdx.x, l = lcnt[s];
  if (l <= 0) return;
  const int o = off[s], n = off[s + 1] - o;
  const double* c = coef + (size_t)s * cld;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double z = 0.0;
    for (int j = 0; j < l; ++j) z += c[j] * V[(size_t)j * ldv + o + i];
    x[o + i] += z;
  }
}

}  // namespace

struct eb_gmres_batch_s {
  int nsys = 0, maxl = 0, ntot = 0, hld = 0;
  hipStream_t st = nullptr;
  std::vector<int> off, maxl_s;
  int* d_off = nullptr;
  int* d_act = nullptr;     // per system: 1 while it iterates (also the re-orth flags)
  int* d_lcnt = nullptr;    // per system: number of Krylov vectors in its correction
  double* d_fac = nullptr;  // per-system scale factors
  double* V = nullptr;      // ntot x (maxl + 1)
  double* r = nullptr;      // residuals (ntot)
  double* hs = nullptr;     // nsys x hld scalars
  double* coef = nullptr;   // nsys x maxl combination coefficients
  // host state
  std::vector<double> hes, q, h, fac, cf;
  std::vector<int> act, flag, lcnt;
  // device-side control (eb_gmres_batch_solve_async)
  GmbState* d_state = nullptr;
  int* d_flag = nullptr;
  int* d_maxl = nullptr;
  int* d_fail = nullptr;
  double* d_hes = nullptr;
  double* d_q = nullptr;
  double* d_cf = nullptr;  // nsys x (maxl + 1): DHELS right-hand sides and coefficients
};

extern "C" int eb_gmres_batch_create(eb_gmres_batch* out, int32_t nsys, const int32_t* off, int32_t maxl,
                                     void* stream) {
  if (!out || nsys < 1 || !off || off[0] != 0) return (eb_set_error("eb_gmres_batch_create: bad arguments"), EB_ERR_ARG);
  *out = nullptr;
  if (maxl == 0) maxl = 10;  // dgmres.f: MAXL = IGWK(1), 0 -> 10
  for (int s = 0; s < nsys; ++s)
    if (off[s + 1] <= off[s]) return (eb_set_error("eb_gmres_batch_create: system %d is empty", s), EB_ERR_ARG);
  eb_gmres_batch g = new eb_gmres_batch_s();
  g->nsys = nsys;
  g->off.assign(off, off + nsys + 1);
  g->ntot = off[nsys];
  g->maxl_s.resize(nsys);
  int mx = 1;
  for (int s = 0; s < nsys; ++s) {
    g->maxl_s[s] = std::max(1, std::min<int>(maxl, off[s + 1] - off[s]));  // MAXL cut to N per system
    mx = std::max(mx, g->maxl_s[s]);
  }
  g->maxl = mx;
  g->hld = mx + 4;
  g->st = (hipStream_t)stream;
  const size_t ns = (size_t)nsys;
  const bool ok = hipMalloc((void**)&g->d_off, sizeof(int) * (ns + 1)) == hipSuccess &&
                  hipMalloc((void**)&g->d_act, sizeof(int) * ns) == hipSuccess &&
                  hipMalloc((void**)&g->d_lcnt, sizeof(int) * ns) == hipSuccess &&
                  hipMalloc((void**)&g->d_fac, sizeof(double) * ns) == hipSuccess &&
                  hipMalloc((void**)&g->V, sizeof(double) * (size_t)g->ntot * (mx + 1)) == hipSuccess &&
                  hipMalloc((void**)&g->r, sizeof(double) * (size_t)g->ntot) == hipSuccess &&
                  hipMalloc((void**)&g->hs, sizeof(double) * ns * g->hld) == hipSuccess &&
                  hipMalloc((void**)&g->coef, sizeof(double) * ns * mx) == hipSuccess &&
                  hipMalloc((void**)&g->d_state, sizeof(GmbState) * ns) == hipSuccess &&
                  hipMalloc((void**)&g->d_flag, sizeof(int) * ns) == hipSuccess &&
                  hipMalloc((void**)&g->d_maxl, sizeof(int) * ns) == hipSuccess &&
                  hipMalloc((void**)&g->d_fail, sizeof(int)) == hipSuccess &&
                  hipMalloc((void**)&g->d_hes, sizeof(double) * ns * (mx + 1) * mx) == hipSuccess &&
                  hipMalloc((void**)&g->d_q, sizeof(double) * ns * 2 * mx) == hipSuccess &&
                  hipMalloc((void**)&g->d_cf, sizeof(double) * ns * (mx + 1)) == hipSuccess &&
                  hipMemcpyAsync(g->d_off, off, sizeof(int) * (ns + 1), hipMemcpyHostToDevice, g->st) == hipSuccess &&
                  hipMemcpyAsync(g->d_maxl, g->maxl_s.data(), sizeof(int) * ns, hipMemcpyHostToDevice, g->st) ==
                      hipSuccess &&
                  hipMemsetAsync(g->d_fail, 0, sizeof(int), g->st) == hipSuccess &&
                  hipStreamSynchronize(g->st) == hipSuccess;
  if (!ok) {
    eb_gmres_batch_destroy(g);
    return (eb_set_error("eb_gmres_batch_create: device allocation failed"), EB_ERR_NOMEM);
  }
  g->hes.assign(ns * (mx + 1) * mx, 0.0);
  g->q.assign(ns * 2 * mx, 0.0);
  g->h.assign(ns * g->hld, 0.0);
  g->fac.assign(ns, 0.0);
  g->cf.assign(ns * mx, 0.0);
  g->act.assign(ns, 0);
  g->flag.assign(ns, 0);
  g->lcnt.assign(ns, 0);
  *out = g;
  return EB_OK;
}

extern "C" int eb_gmres_batch_destroy(eb_gmres_batch g) {
  if (!g) return EB_OK;
  if (g->st) (void)hipStreamSynchronize(g->st);
  void* p[] = {g->d_off, g->d_act, g->d_lcnt, g->d_fac, g->V,      g->r,    g->hs,   g->coef,
               g->d_state, g->d_flag, g->d_maxl, g->d_fail, g->d_hes, g->d_q, g->d_cf};
  for (void* q : p)
    if (q) (void)hipFree(q);
  delete g;
  return EB_OK;
}

extern "C" int eb_gmres_batch_solve(eb_gmres_batch g, eb_apply_fn apply, void* ctx, const double* b, double* x,
                                    double tol, int32_t* iter, double* err, int32_t* ierr) {
  if (!g || !apply || !b || !x) return (eb_set_error("eb_gmres_batch_solve: bad arguments"), EB_ERR_ARG);
  if (tol == 0.0) tol = 500.0 * 1.1102230246251565e-16;  // 500 * D1MACH(3)
  const int S = g->nsys, hld = g->hld, mx = g->maxl, ldh = mx + 1;
  const size_t ldv = (size_t)g->ntot;
  hipStream_t st = g->st;
  std::vector<double> bnrm(S), r0nrm(S), prod(S, 1.0), rho(S), snormw(S, 0.0), e(S, 0.0);
  std::vector<int> it(S, 0), ie(S, 0);
  auto H = [&](int s, int i) -> double& { return g->h[(size_t)s * hld + i]; };
  auto upload_act = [&](const std::vector<int>& a) {
    return hipMemcpyAsync(g->d_act, a.data(), sizeof(int) * S, hipMemcpyHostToDevice, st);
  };
  auto fetch_h = [&]() -> int {
    EB_HIP_CHECK(hipMemcpyAsync(g->h.data(), g->hs, sizeof(double) * (size_t)S * hld, hipMemcpyDeviceToHost, st));
    EB_HIP_CHECK(hipStreamSynchronize(st));
    return EB_OK;
  };
  std::fill(g->hes.begin(), g->hes.end(), 0.0);
  std::fill(g->q.begin(), g->q.end(), 0.0);
  std::fill(g->lcnt.begin(), g->lcnt.end(), 0);
  // r = b - A x, BNRM = |b| per system (every system active for the operator)
  hipLaunchKernelGGL(k_gmb_all_active, dim3((S + 127) / 128), dim3(128), 0, st, S, g->d_act);
  int rc = apply(ctx, x, g->r, st);
  if (rc != EB_OK) return rc;
  hipLaunchKernelGGL(k_gmb_residual, dim3(S), dim3(GB_T), 0, st, g->d_off, b, g->r, g->hs, hld);
  EB_HIP_CHECK(hipGetLastError());
  if ((rc = fetch_h()) != EB_OK) return rc;
  int nact = 0;
  for (int s = 0; s < S; ++s) {
    bnrm[s] = sqrt(H(s, 0));
    r0nrm[s] = sqrt(H(s, 1));
    e[s] = r0nrm[s] / bnrm[s];  // ISDGMR, ITOL = 0, at the start of the cycle
    g->act[s] = e[s] <= tol ? 0 : 1;
    g->fac[s] = g->act[s] ? 1.0 / r0nrm[s] : 0.0;
    nact += g->act[s];
  }
  EB_HIP_CHECK(upload_act(g->act));
  EB_HIP_CHECK(hipMemcpyAsync(g->d_fac, g->fac.data(), sizeof(double) * S, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_gmb_scale, dim3(S), dim3(256), 0, st, g->d_off, g->d_act, g->d_fac, g->r, g->V);
  for (int ll = 1; ll <= mx && nact > 0; ++ll) {
    double* w = g->V + (size_t)ll * ldv;
    if ((rc = apply(ctx, g->V + (size_t)(ll - 1) * ldv, w, st)) != EB_OK) return rc;
    hipLaunchKernelGGL(k_gmb_orth, dim3(S), dim3(GB_T), 0, st, ll, g->d_off, g->d_act, g->V, ldv, g->hs, hld, mx);
    EB_HIP_CHECK(hipGetLastError());
    if ((rc = fetch_h()) != EB_OK) return rc;
    int nflag = 0;
    for (int s = 0; s < S; ++s) {
      g->flag[s] = 0;
      if (!g->act[s]) continue;
      const double vnrm = sqrt(H(s, mx));
      snormw[s] = sqrt(H(s, mx + 1));
      if (!(vnrm + 0.001 * snormw[s] != vnrm)) {
        g->flag[s] = 1;
        ++nflag;
      }
    }
    if (nflag) {
      EB_HIP_CHECK(upload_act(g->flag));
      hipLaunchKernelGGL(k_gmb_reorth, dim3(S), dim3(GB_T), 0, st, ll, g->d_off, g->d_act, g->V, ldv, g->hs, hld, mx);
      EB_HIP_CHECK(hipGetLastError());
      if ((rc = fetch_h()) != EB_OK) return rc;
      for (int s = 0; s < S; ++s)
        if (g->flag[s]) {
          const double sumdsq = H(s, mx + 2);
          if (sumdsq != 0.0) snormw[s] = sqrt(std::max(0.0, snormw[s] * snormw[s] - sumdsq));
        }
    }
    for (int s = 0; s < S; ++s) {
      g->fac[s] = 0.0;
      if (!g->act[s]) continue;
      double* hes = g->hes.data() + (size_t)s * ldh * mx;
      double* qs = g->q.data() + (size_t)s * 2 * mx;
      for (int i = 1; i <= ll; ++i) hes[(size_t)(ll - 1) * ldh + (i - 1)] = H(s, i - 1);
      hes[(size_t)(ll - 1) * ldh + ll] = snormw[s];  // HES(LL+1, LL)
      const int info = dheqr_column(hes, ldh, ll, qs);
      it[s] = ll;
      int done = 0;  // 1: update x with l = ll vectors (converged or out of vectors); 2: stop, no update
      if (info == ll) {
        done = 2;  // singular Hessenberg (dpigmr.f:366, 419-423)
        ie[s] = 2;
      } else {
        prod[s] = prod[s] * qs[2 * ll - 1];
        rho[s] = fabs(prod[s] * r0nrm[s]);
        e[s] = rho[s] / bnrm[s];
        if (e[s] <= tol) {
          done = 1;
          ie[s] = 0;
        } else if (ll == g->maxl_s[s]) {
          if (!(rho[s] < r0nrm[s])) {
            done = 2;
            ie[s] = 2;
          } else {
            done = 1;
            ie[s] = 1;  // IGWK(5) = -1: no restart
          }
        }
      }
      if (done) {
        g->act[s] = 0;
        --nact;
        if (done == 1) {
          std::vector<double> bb(ll + 1, 0.0);
          bb[0] = r0nrm[s];
          dhels(hes, ldh, ll, qs, bb.data());
          for (int j = 0; j < ll; ++j) g->cf[(size_t)s * mx + j] = bb[j];
          g->lcnt[s] = ll;
        }
      } else {
        g->fac[s] = 1.0 / snormw[s];
      }
    }
    if (nact > 0) {
      EB_HIP_CHECK(upload_act(g->act));
      EB_HIP_CHECK(hipMemcpyAsync(g->d_fac, g->fac.data(), sizeof(double) * S, hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(k_gmb_scale, dim3(S), dim3(256), 0, st, g->d_off, g->d_act, g->d_fac, w, w);
    }
  }
  EB_HIP_CHECK(hipMemcpyAsync(g->d_lcnt, g->lcnt.data(), sizeof(int) * S, hipMemcpyHostToDevice, st));
  EB_HIP_CHECK(hipMemcpyAsync(g->coef, g->cf.data(), sizeof(double) * (size_t)S * mx, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_gmb_combine, dim3(S), dim3(256), 0, st, g->d_off, g->d_lcnt, g->coef, mx, g->V, ldv, x);
  EB_HIP_CHECK(hipGetLastError());
  EB_HIP_CHECK(hipStreamSynchronize(st));
  for (int s = 0; s < S; ++s) {
    if (iter) iter[s] = it[s];
    if (err) err[s] = e[s];
    if (ierr) ierr[s] = ie[s];
  }
  return EB_OK;
}

extern "C" int eb_gmres_batch_solve_async(eb_gmres_batch g, eb_apply_fn apply, void* ctx, const double* b, double* x,
                                          double tol, int32_t budget) {
  if (!g || !apply || !b || !x || budget < 1)
    return (eb_set_error("eb_gmres_batch_solve_async: bad arguments"), EB_ERR_ARG);
  if (tol == 0.0) tol = 500.0 * 1.1102230246251565e-16;  // 500 * D1MACH(3)
  const int S = g->nsys, hld = g->hld, mx = g->maxl;
  const int nl = std::min<int>(budget, mx);
  const size_t ldv = (size_t)g->ntot;
  hipStream_t st = g->st;
  const dim3 cb((S + 127) / 128), ct(128);
  hipLaunchKernelGGL(k_gmb_all_active, cb, ct, 0, st, S, g->d_act);
  int rc = apply(ctx, x, g->r, st);
  if (rc != EB_OK) return rc;
  hipLaunchKernelGGL(k_gmb_residual, dim3(S), dim3(GB_T), 0, st, g->d_off, b, g->r, g->hs, hld);
  hipLaunchKernelGGL(k_gmb_dstart, cb, ct, 0, st, S, tol, g->hs, hld, g->d_act, g->d_fac, g->d_lcnt, g->d_state);
  hipLaunchKernelGGL(k_gmb_scale, dim3(S), dim3(256), 0, st, g->d_off, g->d_act, g->d_fac, g->r, g->V);
  for (int ll = 1; ll <= nl; ++ll) {
    double* w = g->V + (size_t)ll * ldv;
    if ((rc = apply(ctx, g->V + (size_t)(ll - 1) * ldv, w, st)) != EB_OK) return rc;
    hipLaunchKernelGGL(k_gmb_step, dim3(S), dim3(GB_T), 0, st, ll, tol, nl, g->d_off, g->V, ldv, g->hs, hld, mx,
                       g->d_maxl, g->d_act, g->d_flag, g->d_fac, g->d_lcnt, g->d_cf, mx + 1, g->d_hes, g->d_q,
                       g->d_state, g->d_fail);
  }
  hipLaunchKernelGGL(k_gmb_combine, dim3(S), dim3(256), 0, st, g->d_off, g->d_lcnt, g->d_cf, mx + 1, g->V, ldv, x);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" const int32_t* eb_gmres_batch_active(eb_gmres_batch g) { return g ? g->d_act : nullptr; }

extern "C" int eb_gmres_batch_clear_fail(eb_gmres_batch g) {
  if (!g) return (eb_set_error("eb_gmres_batch_clear_fail: bad arguments"), EB_ERR_ARG);
  hipLaunchKernelGGL(k_gmb_zero, dim3(1), dim3(64), 0, g->st, g->d_fail);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" int eb_gmres_batch_results(eb_gmres_batch g, int32_t* iter, double* err, int32_t* ierr, int32_t* nfail) {
  if (!g) return (eb_set_error("eb_gmres_batch_results: bad arguments"), EB_ERR_ARG);
  std::vector<GmbState> h(g->nsys);
  int f = 0;
  EB_HIP_CHECK(hipMemcpyAsync(h.data(), g->d_state, sizeof(GmbState) * g->nsys, hipMemcpyDeviceToHost, g->st));
  EB_HIP_CHECK(hipMemcpyAsync(&f, g->d_fail, sizeof(int), hipMemcpyDeviceToHost, g->st));
  EB_HIP_CHECK(hipStreamSynchronize(g->st));
  for (int s = 0; s < g->nsys; ++s) {
    if (iter) iter[s] = h[s].it;
    if (err) err[s] = h[s].err;
    if (ierr) ierr[s] = h[s].ie;
  }
  if (nfail) *nfail = f;
  return EB_OK;
}

extern "C" int eb_sweep_solve_async(eb_sweep sw, eb_gmres_batch g, const double* rhs, double* soln, int32_t budget,
                                    void* stream) {
  if (!sw || !g || !rhs || !soln) return (eb_set_error("eb_sweep_solve_async: bad arguments"), EB_ERR_ARG);
  hipStream_t st = (hipStream_t)stream;
  if (st != g->st)
    return (eb_set_error("eb_sweep_solve_async: the GMRES workspace was created on another stream"), EB_ERR_ARG);
  hipLaunchKernelGGL(k_gm_copy, dim3(ew_grid(g->ntot)), dim3(256), 0, st, g->ntot, rhs, soln);
  eb_sweep_set_mask(sw, g->d_act);
  const int rc = eb_gmres_batch_solve_async(g, eb_sweep_gmres_apply, sw, rhs, soln, 1.0e-10, budget);
  eb_sweep_set_mask(sw, nullptr);
  return rc;
}

extern "C" int eb_sweep_gmres_apply(void* sw, const double* x, double* y, void* stream) {
  return eb_sweep_apply((eb_sweep)sw, x, y, stream);
}

extern "C" int eb_sweep_solve(eb_sweep sw, eb_gmres_batch g, const double* rhs, double* soln, double* za, double* pk,
                              int32_t* iter, double* err, int32_t* ierr, void* stream) {
  if (!sw || !g || !rhs || !soln) return (eb_set_error("eb_sweep_solve: bad arguments"), EB_ERR_ARG);
  hipStream_t st = (hipStream_t)stream;
  if (st != g->st) return (eb_set_error("eb_sweep_solve: the GMRES workspace was created on another stream"), EB_ERR_ARG);
  const int32_t* uo = eb_sweep_unknown_offsets(sw);
  if ((int)g->off.size() != g->nsys + 1 || !std::equal(g->off.begin(), g->off.end(), uo))
    return (eb_set_error("eb_sweep_solve: the GMRES workspace does not match the sweep's systems"), EB_ERR_ARG);
  // soln = rhs as the initial guesses (stokesbubb.f:343-346), ITOL 0, TOL 1e-10, no restart
  hipLaunchKernelGGL(k_gm_copy, dim3(ew_grid(g->ntot)), dim3(256), 0, st, g->ntot, rhs, soln);
  eb_sweep_set_mask(sw, g->d_act);
  int rc = eb_gmres_batch_solve(g, eb_sweep_gmres_apply, sw, rhs, soln, 1.0e-10, iter, err, ierr);
  eb_sweep_set_mask(sw, nullptr);
  if (rc != EB_OK) return rc;
  if (za || pk) rc = eb_sweep_parts(sw, soln, za, pk, stream);
  return rc;
}
