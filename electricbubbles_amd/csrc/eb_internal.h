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
// The radix sort's tile: its first pass counts the low 8 key bits of keys [RS_TILE_KEYS b,
// RS_TILE_KEYS (b + 1)) into tmp[d * nblk + b] (nblk tiles).  hist0_ready: the caller has
// already written those counts (the FMM's k_keys), so the first pass skips its histogram.
constexpr int RS_TILE_KEYS = 4096;
int radix_sort_pairs(uint32_t* k0, int32_t* v0, uint32_t* k1, int32_t* v1, int n, int nbits, int* tmp,
                     hipStream_t st, bool hist0_ready = false);

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
// The same 27 offsets in the order the M2L kernels take them: the 12 of ring 2 with
// min(|dx|, |dy|) <= 1 first (centre distance |d| = 2 or sqrt 5 box widths), then the 15
// corners of ring 2 and boxes of ring 3 (|d| >= 2 sqrt 2), each group in MKINTL order.  The
// convergence ratio of a translation between scaled expansions is at most sqrt(2) / |d|, so
// the second group reaches the first group's truncation error with fewer input terms; the
// kernels translate it with about two thirds of them (M2L_NEAR sources first).
constexpr int M2L_NEAR = 12;
__host__ __device__ inline int m2l_offsets(int px, int py, int* dxs, int* dys) {
  int ax[27], ay[27];
  const int no = intl_offsets(px, py, ax, ay);
  int n = 0;
  for (int pass = 0; pass < 2; ++pass)
    for (int o = 0; o < no; ++o) {
      const int mx = ax[o] < 0 ? -ax[o] : ax[o], my = ay[o] < 0 ? -ay[o] : ay[o];
      const bool near = (mx > my ? mx : my) == 2 && (mx < my ? mx : my) <= 1;
      if (near == (pass == 0)) {
        dxs[n] = ax[o];
        dys[n] = ay[o];
        ++n;
      }
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
