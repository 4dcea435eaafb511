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

// w -= h v_i (h read from device memory); then vnext . w, or |w|^2 when vnext is NULL
__global__ __launch_bounds__(GM_T) void k_gm_axpy_dot(int n, double* __restrict__ w, const double* __restrict__ vi,
                                                      const double* __restrict__ h, const double* __restrict__ vnext,
                                                      double* part, unsigned* count, double* out) {
  const double c = -*h;
  double s = 0.0;
  for (int i = blockIdx.x * GM_T + threadIdx.x; i < n; i += gridDim.x * GM_T) {
    const double wi = w[i] + c * vi[i];
    w[i] = wi;
    s += vnext ? vnext[i] * wi : wi * wi;
  }
  finish2(s, 0.0, part, count, out, nullptr);
}

// v . w
__global__ __launch_bounds__(GM_T) void k_gm_dot(int n, const double* __restrict__ v, const double* __restrict__ w,
                                                 double* part, unsigned* count, double* out) {
  double s = 0.0;
  for (int i = blockIdx.x * GM_T + threadIdx.x; i < n; i += gridDim.x * GM_T) s += v[i] * w[i];
  finish2(s, 0.0, part, count, out, nullptr);
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
int dheqr_column(std::vector<double>& hes, int ldh, int ll, std::vector<double>& q) {
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
void dhels(const std::vector<double>& hes, int ldh, int n, const std::vector<double>& q, std::vector<double>& b) {
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

int fetch(eb_gmres_s* g, int first, int count, double* host) {
  EB_HIP_CHECK(hipMemcpyAsync(host, g->scal + first, sizeof(double) * count, hipMemcpyDeviceToHost, g->st));
  EB_HIP_CHECK(hipStreamSynchronize(g->st));
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
    hipLaunchKernelGGL(k_gm_norm_dot, dim3(nb), dim3(GM_T), 0, st, n, w, col(g, 0), g->red.part, g->red.count,
                       g->scal + s_vn(g), g->scal + S_H + 0);
    for (int i = 1; i <= ll; ++i) {
      const double* vnext = i < ll ? col(g, i) : nullptr;
      double* out = i < ll ? g->scal + S_H + i : g->scal + s_sn(g);
      hipLaunchKernelGGL(k_gm_axpy_dot, dim3(nb), dim3(GM_T), 0, st, n, w, col(g, i - 1), g->scal + S_H + (i - 1),
                         vnext, g->red.part, g->red.count, out);
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
    const int info = dheqr_column(hes, ldh, ll, q);
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
  dhels(hes, ldh, l, q, b);
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
            hipMalloc((void**)&g->red.count, sizeof(unsigned)) == hipSuccess;
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
  void* p[] = {g->V, g->r, g->scal, g->coef, g->red.part, g->red.count};
  for (void* q : p)
    if (q) (void)hipFree(q);
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
