// eb_fft.h -- the per-body FFT of the spectral kernels: one workgroup transforms one body's
// sequence in LDS, in place.
//
// A length L = R1 R2 ... Rs is transformed by Stockham stages.  Stage (R, Ns), Ns = R1..R(i-1),
// has L/R butterflies; butterfly t (k = t mod Ns, j = t / Ns) reads x[t + r L/R], multiplies
// input r by W_{Ns R}^{k r}, forms the R-point DFT and writes output q to
// x[j Ns R + k + q Ns].  The result is in natural order after the last stage.  Every thread
// first loads all of its butterflies into registers, the workgroup synchronises, and then
// every thread writes: the stage runs in place, so a transform needs no scratch sequence.
//
// Stage kinds:
//  * radix 2, 3, 4, 5, 8: closed-form butterflies;
//  * any other prime R (7, 11, 13, ... and large primes that do not fit the Bluestein
//    workspace): each thread forms whole outputs as R-term sums with twiddles from the
//    table of L-th roots of unity;
//  * the largest prime factor P >= 17, when it fits: the last stage becomes Ns = L / P
//    independent length-P DFTs of pre-twiddled inputs, each done by Bluestein's chirp-z
//    method (W^{rq} = c_r c_q conj(c_{q-r}), c_n = exp(-i pi n^2 / P)) as a circular
//    convolution of length M >= 2P - 1, M = 2^a 3^b 5^c, through two length-M transforms of
//    this same engine, batched over the Ns sub-sequences, in an LDS workspace of Ns M entries.
//
// The reference's FFTPACK (dcfft.f) uses the same DFT definition (forward exp(-2 pi i jk/N),
// unnormalised; backward exp(+2 pi i jk/N)); results agree to rounding.
//
// The per-element index functions are __host__ __device__, so the stage sequencing can be
// checked on the host by a sequential emulation of the workgroup.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ebfft {

constexpr int FFT_THREADS = 512;  // threads of the spectral workgroups
constexpr int EMAX = 8192;        // largest batched length (Ns M) a stage handles in registers
constexpr int LMAX = 3200;        // largest body length
constexpr int BL_MIN = 17;        // smallest prime done by Bluestein

struct Stage {
  int32_t R;    // radix (2, 3, 4, 5, 8: closed form; other: direct sum)
  int32_t Ns;   // product of the previous radices
  int32_t LR;   // L / R
  int32_t tw;   // [r-1][k] table of W_{Ns R}^{k r}, r = 1..R-1, k < Ns (closed-form radices)
  uint32_t mNs; // magic multipliers: t / Ns = umulhi(t, mNs), t / LR = umulhi(t, mLR)
  uint32_t mLR;
};

struct Plan {
  int32_t L;       // transform length
  int32_t st, nst; // stages [st, st + nst) of the stage array (before the Bluestein stage)
  int32_t root;    // table offset of W_L^j, j < L (direct-sum stages)
  // Bluestein last stage (P = 0: none)
  int32_t P, Ns, M;
  int32_t pre;     // [s][r]: W_L^{r s} c_r, s < Ns, r < P
  int32_t post;    // c_q, q < P
  int32_t H;       // DFT_M of the chirp filter, divided by M
  int32_t mst, mnst;  // the stages of the length-M transform
  uint32_t mM, mNsB;  // magic for / M and / Ns
  int32_t cost;    // relative run time, for the body order
};

// ---- complex arithmetic usable on host and device ----------------------------------------
__host__ __device__ inline double2 c2(double a, double b) { return make_double2(a, b); }
__host__ __device__ inline double2 cmulx(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__host__ __device__ inline double2 caddx(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csubx(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// s i z for s = +-1
__host__ __device__ inline double2 cmuli(double2 z, double s) { return make_double2(-s * z.y, s * z.x); }
__host__ __device__ inline double2 conjif(double2 z, int sign) { return sign > 0 ? make_double2(z.x, -z.y) : z; }

// t / d for t * d < 2^32, with m = floor(2^32 / d) + 1 (d >= 2) or m = 0 (d = 1)
__host__ __device__ inline int fdivm(int t, int d, uint32_t m) {
#ifdef __HIP_DEVICE_COMPILE__
  return d == 1 ? t : (int)__umulhi((uint32_t)t, m);
#else
  return d == 1 ? t : (int)(((uint64_t)(uint32_t)t * m) >> 32);
#endif
}
inline uint32_t magic_for(int d) { return d <= 1 ? 0u : (uint32_t)((1ull << 32) / (uint64_t)d + 1ull); }

// ---- closed-form butterflies: v[q] <- sum_r v[r] W^{rq}, W = exp(sign 2 pi i / R) ---------
__host__ __device__ inline void bfly2(double2* v) {
  const double2 a = v[0], b = v[1];
  v[0] = caddx(a, b);
  v[1] = csubx(a, b);
}
__host__ __device__ inline void bfly4(double2* v, double s) {
  const double2 t0 = caddx(v[0], v[2]), t1 = csubx(v[0], v[2]);
  const double2 t2 = caddx(v[1], v[3]), t3 = cmuli(csubx(v[1], v[3]), s);
  v[0] = caddx(t0, t2);
  v[2] = csubx(t0, t2);
  v[1] = caddx(t1, t3);
  v[3] = csubx(t1, t3);
}
__host__ __device__ inline void bfly3(double2* v, double s) {
  const double h = 0.86602540378443864676372317075293618;  // sin(2 pi / 3)
  const double2 t = caddx(v[1], v[2]);
  const double2 d = cmuli(csubx(v[1], v[2]), s * h);
  const double2 m = make_double2(v[0].x - 0.5 * t.x, v[0].y - 0.5 * t.y);
  v[0] = caddx(v[0], t);
  v[1] = caddx(m, d);
  v[2] = csubx(m, d);
}
__host__ __device__ inline void bfly5(double2* v, double s) {
  const double c1 = 0.30901699437494742410229341718281906;   // cos(2 pi / 5)
  const double c2 = -0.80901699437494742410229341718281906;  // cos(4 pi / 5)
  const double s1 = 0.95105651629515357211643933337938214;   // sin(2 pi / 5)
  const double s2 = 0.58778525229247312916870595463907277;   // sin(4 pi / 5)
  const double2 a0 = v[0];
  const double2 t1 = caddx(v[1], v[4]), t2 = caddx(v[2], v[3]);
  const double2 d1 = csubx(v[1], v[4]), d2 = csubx(v[2], v[3]);
  const double2 m1 = make_double2(a0.x + c1 * t1.x + c2 * t2.x, a0.y + c1 * t1.y + c2 * t2.y);
  const double2 m2 = make_double2(a0.x + c2 * t1.x + c1 * t2.x, a0.y + c2 * t1.y + c1 * t2.y);
  const double2 e1 = cmuli(make_double2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y), s);
  const double2 e2 = cmuli(make_double2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y), s);
  v[0] = caddx(a0, caddx(t1, t2));
  v[1] = caddx(m1, e1);
  v[4] = csubx(m1, e1);
  v[2] = caddx(m2, e2);
  v[3] = csubx(m2, e2);
}
// radix 8 as 2 x 4: u_r = a_r + a_{r+4} gives the even outputs, (a_r - a_{r+4}) W^r the odd
__host__ __device__ inline void bfly8(double2* v, double s) {
  const double h = 0.70710678118654752440084436210484904;  // 1 / sqrt 2
  double2 u[4], w[4];
  for (int r = 0; r < 4; ++r) {
    u[r] = caddx(v[r], v[r + 4]);
    w[r] = csubx(v[r], v[r + 4]);
  }
  // W^1 = (1 + s i) h, W^2 = s i, W^3 = (-1 + s i) h
  w[1] = make_double2(h * (w[1].x - s * w[1].y), h * (w[1].y + s * w[1].x));
  w[2] = cmuli(w[2], s);
  w[3] = make_double2(h * (-w[3].x - s * w[3].y), h * (-w[3].y + s * w[3].x));
  bfly4(u, s);
  bfly4(w, s);
  for (int q = 0; q < 4; ++q) {
    v[2 * q] = u[q];
    v[2 * q + 1] = w[q];
  }
}
template <int R>
__host__ __device__ inline void bfly(double2* v, double s) {
  if constexpr (R == 2) bfly2(v);
  else if constexpr (R == 3) bfly3(v, s);
  else if constexpr (R == 4) bfly4(v, s);
  else if constexpr (R == 5) bfly5(v, s);
  else bfly8(v, s);
}

// ---- closed-form stage over nseq sequences of length L stored back to back ----------------
// butterfly t < nseq L / R: its sequence b = t / LR and position tt = t - b LR
template <int R>
__host__ __device__ inline void spec_load(const double2* x, int L, int t, const Stage& s, const double2* tw, int sign,
                                          double2* v) {
  const int b = fdivm(t, s.LR, s.mLR);
  const int tt = t - b * s.LR;
  const double2* src = x + b * L + tt;
  v[0] = src[0];
  if (s.Ns == 1) {
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = src[r * s.LR];
  } else {
    const int k = tt - fdivm(tt, s.Ns, s.mNs) * s.Ns;
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmulx(src[r * s.LR], conjif(tw[(r - 1) * s.Ns + k], sign));
  }
}
template <int R>
__host__ __device__ inline void spec_store(double2* x, int L, int t, const Stage& s, const double2* v) {
  const int b = fdivm(t, s.LR, s.mLR);
  const int tt = t - b * s.LR;
  const int j = fdivm(tt, s.Ns, s.mNs), k = tt - j * s.Ns;
  double2* dst = x + b * L + j * s.Ns * R + k;
#pragma unroll
  for (int q = 0; q < R; ++q) dst[q * s.Ns] = v[q];
}

// ---- direct-sum stage, one sequence: output o = j Ns R + q Ns + k -------------------------
__host__ __device__ inline double2 gen_out(const double2* x, int L, int o, const Stage& s, const double2* root,
                                           int sign) {
  const int R = s.R, Ns = s.Ns;
  const int k = o % Ns;
  const int q = (o / Ns) % R;
  const int j = o / (Ns * R);
  const int t = j * Ns + k;
  const int e = (k + q * Ns) * (L / (Ns * R));  // W_{Ns R}^{r (k + q Ns)} = W_L^{r e}
  double2 acc = x[t];
  int idx = e;
  for (int r = 1; r < R; ++r) {
    const double2 w = conjif(root[idx], sign);
    const double2 a = x[t + r * s.LR];
    acc.x = fma(a.x, w.x, fma(-a.y, w.y, acc.x));
    acc.y = fma(a.x, w.y, fma(a.y, w.x, acc.y));
    idx += e;
    if (idx >= L) idx -= L;
  }
  return acc;
}

// ---- Bluestein stage ---------------------------------------------------------------------
// workspace element e = s M + m: the pre-twiddled, chirped input of sub-sequence s (0 past P)
__host__ __device__ inline double2 blue_in(const double2* x, int e, const Plan& p, const double2* tab, int sign) {
  const int s = fdivm(e, p.M, p.mM);
  const int m = e - s * p.M;
  if (m >= p.P) return make_double2(0.0, 0.0);
  return cmulx(conjif(x[s + m * p.Ns], sign), tab[p.pre + s * p.P + m]);
}
// output o = q Ns + s of the stage
__host__ __device__ inline double2 blue_out(const double2* U, int o, const Plan& p, const double2* tab, int sign) {
  const int q = fdivm(o, p.Ns, p.mNsB);
  const int s = o - q * p.Ns;
  return conjif(cmulx(tab[p.post + q], U[s * p.M + q]), sign);
}
__host__ __device__ inline double2 blue_mul(const double2* U, int e, const Plan& p, const double2* tab) {
  const int s = fdivm(e, p.M, p.mM);
  return cmulx(U[e], tab[p.H + (e - s * p.M)]);
}

// ---- device driver -----------------------------------------------------------------------
#ifdef __HIPCC__
template <int R>
__device__ __forceinline__ void run_spec(double2* x, int L, int nseq, const Stage& s, const double2* tab, int sign) {
  constexpr int NB = (EMAX / R + FFT_THREADS - 1) / FFT_THREADS;
  const int nbf = nseq * s.LR;
  const double2* tw = tab + s.tw;
  double2 v[NB][R];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int t = (int)threadIdx.x + i * FFT_THREADS;
    if (t < nbf) spec_load<R>(x, L, t, s, tw, sign, v[i]);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const int t = (int)threadIdx.x + i * FFT_THREADS;
    if (t < nbf) {
      bfly<R>(v[i], (double)sign);
      spec_store<R>(x, L, t, s, v[i]);
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void run_gen(double2* x, int L, const Stage& s, const double2* root, int sign) {
  constexpr int NO = (LMAX + FFT_THREADS - 1) / FFT_THREADS;
  double2 y[NO];
#pragma unroll
  for (int i = 0; i < NO; ++i) {
    const int o = (int)threadIdx.x + i * FFT_THREADS;
    if (o < L) y[i] = gen_out(x, L, o, s, root, sign);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NO; ++i) {
    const int o = (int)threadIdx.x + i * FFT_THREADS;
    if (o < L) x[o] = y[i];
  }
  __syncthreads();
}

// stages [st, st + nst) on nseq sequences of length L (closed-form radices and, for one
// sequence, direct sums)
__device__ __forceinline__ void run_stages(double2* x, int L, int nseq, const Stage* stages, int st, int nst, int root,
                                  const double2* tab, int sign) {
  for (int i = 0; i < nst; ++i) {
    const Stage s = stages[st + i];
    switch (s.R) {
      case 2: run_spec<2>(x, L, nseq, s, tab, sign); break;
      case 3: run_spec<3>(x, L, nseq, s, tab, sign); break;
      case 4: run_spec<4>(x, L, nseq, s, tab, sign); break;
      case 5: run_spec<5>(x, L, nseq, s, tab, sign); break;
      case 8: run_spec<8>(x, L, nseq, s, tab, sign); break;
      default: run_gen(x, L, s, tab + root, sign); break;
    }
  }
}

// In-place transform of x[0..L) in LDS by the whole workgroup; U is the Bluestein workspace
// (Ns M entries of LDS, unused without a Bluestein stage).  Starts and ends synchronised
// (the caller synchronises after writing x).
__device__ __forceinline__ void fft_inplace(double2* x, double2* U, const Plan& p, const Stage* stages, const double2* tab,
                                   int sign) {
  run_stages(x, p.L, 1, stages, p.st, p.nst, p.root, tab, sign);
  if (p.P == 0) return;
  const int E = p.Ns * p.M;
  for (int e = threadIdx.x; e < E; e += FFT_THREADS) U[e] = blue_in(x, e, p, tab, sign);
  __syncthreads();
  run_stages(U, p.M, p.Ns, stages, p.mst, p.mnst, 0, tab, -1);
  for (int e = threadIdx.x; e < E; e += FFT_THREADS) U[e] = blue_mul(U, e, p, tab);
  __syncthreads();
  run_stages(U, p.M, p.Ns, stages, p.mst, p.mnst, 0, tab, +1);
  for (int o = threadIdx.x; o < p.L; o += FFT_THREADS) x[o] = blue_out(U, o, p, tab, sign);
  __syncthreads();
}
#endif

}  // namespace ebfft
