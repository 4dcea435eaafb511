// eb_internal.h -- declarations shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/eb_cadap.h"
#include "eb_common.h"

extern "C" void eb_set_error(const char* fmt, ...);

namespace eb {
size_t scan_temp_ints(int n);
int scan_exclusive(const int* in, int* out, int n, bool write_total, int* tmp, hipStream_t st);
size_t radix_sort_temp_ints(int n);
int radix_sort_pairs(uint32_t* k0, int32_t* v0, uint32_t* k1, int32_t* v1, int n, int nbits, int* tmp,
                     hipStream_t st);

// MKINTL offsets (cadap2.f:800-827) for a target whose 0-based column/row parity is
// (px, py): column-outer, row-inner, skipping the 3x3 neighbourhood.  Returns 27.
__host__ __device__ inline int intl_offsets(int px, int py, int* dxs, int* dys) {
  // 1-based column even  <=>  0-based px = 1  ->  columns icol-3 .. icol+2
  const int dx0 = px ? -3 : -2, dy0 = py ? -3 : -2;
  int n = 0;
  for (int dx = dx0; dx <= dx0 + 5; ++dx)
    for (int dy = dy0; dy <= dy0 + 5; ++dy) {
      if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1) continue;
      dxs[n] = dx;
      dys[n] = dy;
      ++n;
    }
  return n;
}
// Depth below the segments' boxes for a segmented plan whose largest segment has nmax points
// spread along curves: the number of leaves across a segment (2^d) grows until the curve
// leaves hold about 40 points (nmax / 2^d <= 128, about pi leaves per box side crossed), within
// the plan's 10 levels.
inline int seg_levmax(int nmax, int nseg) {
  int k = 0;
  while ((1 << (2 * k)) < nseg) ++k;
  int d = 2;
  while (d < 10 - k && (nmax >> d) > 128) ++d;
  return d;
}
}  // namespace eb
