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
