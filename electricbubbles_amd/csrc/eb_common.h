// eb_common.h -- shared device helpers for the ElectricBubbles MI355X kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define EB_WAVE 64

namespace eb {

// ---- complex helpers on double2 (re, im) -------------------------------------------------
__device__ __forceinline__ double2 cmk(double a, double b) { return make_double2(a, b); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// acc += a*b
__device__ __forceinline__ void cfma(double2& acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
  acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
}
// a*b + c in four FMAs (cadd(cmul(a, b), c) compiles to two multiplies, two FMAs and two adds)
__device__ __forceinline__ double2 cmuladd(double2 a, double2 b, double2 c) {
  return make_double2(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}
// acc += s*a  (s real)
__device__ __forceinline__ void rfma(double2& acc, double s, double2 a) {
  acc.x = fma(s, a.x, acc.x);
  acc.y = fma(s, a.y, acc.y);
}

// 1/a for a > 0: hardware reciprocal estimate refined by two Newton-Raphson steps
// y <- y (2 - a y), written as e = 1 - a y ; y = y + y e.
// 1/a from v_rcp_f64 and one Newton step (relative error below 1e-14; the near field's
// contribution to GW needs no more)
__device__ __forceinline__ double rcp_nr1(double a) {
  const double y = __builtin_amdgcn_rcp(a);
  return fma(y, fma(-a, y, 1.0), y);
}

__device__ __forceinline__ double rcp_nr(double a) {
  double y = __builtin_amdgcn_rcp(a);
  double e = fma(-a, y, 1.0);
  y = fma(y, e, y);
  e = fma(-a, y, 1.0);
  y = fma(y, e, y);
  return y;
}

// complex reciprocal 1/z
__device__ __forceinline__ double2 crcp(double2 z) {
  double ir = rcp_nr(fma(z.x, z.x, z.y * z.y));
  return make_double2(z.x * ir, -z.y * ir);
}

// One source-target pair of the near field, in the translation-invariant form
//   qa/d + sgn (conj(qb)/conj(d) - conj(qa) d/conj(d)^2),  d = z_src - z_tgt
// (SLDIRDR / SLSELFDR, cadapdrop2.f:586-599, 741-767, with QC = -conj(z) QA folded in).
// acc1 collects qa/d, acc2 the bracket.  With MASK, `self` (source == target) contributes
// exactly zero: the reciprocal is forced to 0, which zeroes both terms without a branch.
// FAST uses one Newton step on the hardware reciprocal, the exact variant two.
template <bool MASK, bool FAST>
__device__ __forceinline__ void p2p_pair(double2 zi, double2 zj, double2 qa, double2 qb, bool self, double2& acc1,
                                         double2& acc2) {
  const double dx = zj.x - zi.x, dy = zj.y - zi.y;
  const double r2 = fma(dx, dx, dy * dy);
  double ir;
  if (MASK) {
    ir = FAST ? rcp_nr1(self ? 1.0 : r2) : rcp_nr(self ? 1.0 : r2);
    ir = self ? 0.0 : ir;
  } else {
    ir = FAST ? rcp_nr1(r2) : rcp_nr(r2);
  }
  const double wx = dx * ir, wy = dy * ir;  // w = 1/conj(d)
  // acc1 += qa * conj(w)
  acc1.x = fma(qa.x, wx, fma(qa.y, wy, acc1.x));
  acc1.y = fma(qa.y, wx, fma(-qa.x, wy, acc1.y));
  // e = d * w = d / conj(d)
  const double ex = fma(dx, wx, -dy * wy);
  const double ey = fma(dx, wy, dy * wx);
  // v = conj(qb) - conj(qa) * e
  const double vx = fma(-qa.x, ex, fma(-qa.y, ey, qb.x));
  const double vy = fma(-qa.x, ey, fma(qa.y, ex, -qb.y));
  // acc2 += v * w
  acc2.x = fma(vx, wx, fma(-vy, wy, acc2.x));
  acc2.y = fma(vx, wy, fma(vy, wx, acc2.y));
}

// ---- Morton helpers (x in the even bits) ---------------------------------------------------
__host__ __device__ __forceinline__ uint32_t spread2(uint32_t v) {
  v &= 0x0000ffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}
__host__ __device__ __forceinline__ uint32_t compact2(uint32_t v) {
  v &= 0x55555555u;
  v = (v | (v >> 1)) & 0x33333333u;
  v = (v | (v >> 2)) & 0x0f0f0f0fu;
  v = (v | (v >> 4)) & 0x00ff00ffu;
  v = (v | (v >> 8)) & 0x0000ffffu;
  return v;
}
__host__ __device__ __forceinline__ uint32_t morton(uint32_t ix, uint32_t iy) { return spread2(ix) | (spread2(iy) << 1); }

}  // namespace eb

#define EB_HIP_CHECK(call)                                                   \
  do {                                                                       \
    hipError_t _e = (call);                                                  \
    if (_e != hipSuccess) {                                                  \
      eb_set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call,        \
                   hipGetErrorString(_e));                                   \
      return EB_ERR_HIP;                                                     \
    }                                                                        \
  } while (0)
