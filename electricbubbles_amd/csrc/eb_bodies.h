// eb_bodies.h -- the per-body machinery shared by the spectral and field kernels: the body
// set (lengths, offsets, FFT plans), its device view, and block-level helpers (FFT in LDS,
// FDIFFF wavenumbers, FILTER, block sums).  One workgroup of SP_THREADS threads owns one body.
#pragma once
#include <math.h>

#include <vector>

#include "../../include/eb_spectral.h"
#include "eb_fft.h"
#include "eb_internal.h"

constexpr int SP_THREADS = ebfft::FFT_THREADS;

struct eb_bodies_s {
  int nbody = 0, n = 0, nmax = 0, ext = 0;
  std::vector<int> nd, off;
  int* d_nd = nullptr;       // [nbody]
  int* d_off = nullptr;      // [nbody + 1]
  int* d_plan_of = nullptr;  // [nbody] plan of the body's length
  int* d_order = nullptr;    // [nbody] bodies by decreasing transform cost
  ebfft::Plan* d_plans = nullptr;
  ebfft::Stage* d_stages = nullptr;
  double2* d_tab = nullptr;  // twiddles, roots, chirps, filter transforms
  double* d_red = nullptr;   // [nbody][4] xmin, xmax, ymin, ymax
};

// ------------------------------------------------------------------------------------------
// per-body helpers in LDS
// ------------------------------------------------------------------------------------------

// wavenumber of coefficient i under FDIFFF's rules (dcfft.f:1404-1422): 0 for the mean and,
// for even N, the Nyquist coefficient
__device__ __forceinline__ double fdiff_k(int i, int N) {
  if (i == 0) return 0.0;
  if ((N & 1) == 0) {
    const int h = N / 2;
    if (i < h) return (double)i;
    if (i == h) return 0.0;
    return (double)(i - N);
  }
  return (i <= (N - 1) / 2) ? (double)i : (double)(i - N);
}

// FILTER (electric.f:1152-1187) on coefficients c[0..N) in LDS: exp(-10 (i/n)^25) with the
// ratio formed in single precision, components below 1e-13 zeroed, Nyquist zeroed.
__device__ __forceinline__ double filt_factor(int i, int n) {
  const double ratio = (double)((float)i / (float)n);
  double r = 1.0, b = ratio;  // ratio**25 by repeated squaring
  for (int e = 25; e > 0; e >>= 1) {
    if (e & 1) r *= b;
    b *= b;
  }
  return exp(-10.0 * r);
}

__device__ __forceinline__ void filter_block(double2* c, int N) {
  const int n = N / 2;
  const double eps = 1e-13;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    int m;
    if (i <= n) m = i;
    else m = N - i;  // i = N - m, m = 1..n-1
    double2 v = c[i];
    if (i == n) {
      v = make_double2(0, 0);
    } else {
      const double f = filt_factor(m, n);
      v.x *= f;
      v.y *= f;
      if (fabs(v.x) < eps) v.x = 0.0;
      if (fabs(v.y) < eps) v.y = 0.0;
    }
    c[i] = v;
  }
  __syncthreads();
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  const double s = red[0];
  __syncthreads();
  return s;
}

struct BodyView {
  const int* nd;
  const int* off;
  const int* plan_of;
  const int* order;
  const ebfft::Plan* plans;
  const ebfft::Stage* stages;
  const double2* tab;
  int nmax;
};

static inline BodyView view(eb_bodies B) {
  return BodyView{B->d_nd, B->d_off, B->d_plan_of, B->d_order, B->d_plans, B->d_stages, B->d_tab, B->nmax};
}

// the body of this workgroup and its transform plan
struct BodyCtx {
  int k, N, o0;
  ebfft::Plan p;
};
__device__ __forceinline__ BodyCtx body_ctx(const BodyView& B) {
  BodyCtx c;
  c.k = B.order[blockIdx.x];
  c.N = B.nd[c.k];
  c.o0 = B.off[c.k];
  c.p = B.plans[B.plan_of[c.k]];
  return c;
}

__device__ __forceinline__ void fft(double2* x, double2* ws, const BodyCtx& c, const BodyView& B, int sign) {
  ebfft::fft_inplace(x, ws, c.p, B.stages, B.tab, sign);
}


template <typename T>
static inline int dalloc(T** p, size_t count) {
  *p = nullptr;
  if (hipMalloc((void**)p, sizeof(T) * (count ? count : 1)) != hipSuccess) {
    eb_set_error("hipMalloc(%zu bytes) failed", sizeof(T) * count);
    return EB_ERR_NOMEM;
  }
  return EB_OK;
}

template <typename T>
static inline bool upload(T* dst, const std::vector<T>& src, hipStream_t st) {
  return src.empty() || hipMemcpyAsync(dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice, st) == hipSuccess;
}

// dynamic LDS of a spectral kernel: nseq sequences of nmax entries and the Bluestein workspace
static inline size_t lds_bytes(eb_bodies B, int nseq) { return sizeof(double2) * ((size_t)B->nmax * nseq + (size_t)B->ext); }
