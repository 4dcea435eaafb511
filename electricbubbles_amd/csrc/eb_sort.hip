// eb_sort.hip -- stable LSD radix sort of (uint32 key, int32 value) pairs and a single-block
// exclusive scan, both hand-written for gfx950 (wave64).  Used by the tree build: stability
// is what makes the per-box source order equal to the reference's source-index order
// (ASSIGN1, cadap2.f:346-353).
#include "eb_internal.h"

namespace eb {

constexpr int RS_THREADS = 256;                 // k_rs_prefix: one thread per digit
constexpr int RS_BT = 1024;                     // k_rs_hist and k_rs_scatter
constexpr int RS_TILE = 4096;                   // elements per tile (one block of either)
constexpr int RS_BW = RS_BT / EB_WAVE;          // 16 waves per block
constexpr int RS_SUB = RS_TILE / RS_BW;         // 256 consecutive elements per wave
constexpr int RS_ROUNDS = RS_SUB / EB_WAVE;     // 4 rounds of 64
constexpr int RS_BITS = 8;
constexpr int RS_RADIX = 1 << RS_BITS;
static_assert(RS_RADIX == RS_THREADS && RS_RADIX <= RS_BT, "digit-indexed threads");
static_assert(RS_TILE == RS_TILE_KEYS, "the tile that callers count ahead of the sort");

// per tile: the count of each digit
__global__ __launch_bounds__(RS_BT) void k_rs_hist(const uint32_t* __restrict__ keys, int n, int shift, int nblk,
                                                   int* __restrict__ hist) {
  __shared__ int h[RS_RADIX];
  if (threadIdx.x < RS_RADIX) h[threadIdx.x] = 0;
  __syncthreads();
  const int base = blockIdx.x * RS_TILE;
#pragma unroll
  for (int r = 0; r < RS_TILE / RS_BT; ++r) {
    const int e = base + r * RS_BT + threadIdx.x;
    if (e < n) atomicAdd(&h[(keys[e] >> shift) & (RS_RADIX - 1)], 1);
  }
  __syncthreads();
  if (threadIdx.x < RS_RADIX) hist[threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// Per digit d (one block each): the exclusive prefix over the tiles of d's tile counts
// (hist[d * nblk + b] -> pre[d * nblk + b]) and d's total over all tiles (tot[d]).  The
// scatter adds the digits' own prefix, so the three launches of a general scan of the
// RS_RADIX x nblk counts become this one.
__global__ __launch_bounds__(RS_THREADS) void k_rs_prefix(const int* __restrict__ hist, int nblk,
                                                          int* __restrict__ pre, int* __restrict__ tot) {
  __shared__ int part[RS_THREADS];
  const int d = blockIdx.x, t = threadIdx.x;
  const int chunk = (nblk + RS_THREADS - 1) / RS_THREADS;
  const int s0 = min(nblk, t * chunk), e0 = min(nblk, s0 + chunk);
  const int* h = hist + (size_t)d * nblk;
  int sum = 0;
  for (int b = s0; b < e0; ++b) sum += h[b];
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < RS_THREADS; off <<= 1) {
    const int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - sum;
  int* o = pre + (size_t)d * nblk;
  for (int b = s0; b < e0; ++b) {
    const int x = h[b];
    o[b] = run;
    run += x;
  }
  if (t == RS_THREADS - 1) tot[d] = part[t];
}

// Exclusive scan of v over threads 0..RS_RADIX-1 of the block (the others pass 0 and get 0):
// wave scans by shuffles, then the wave totals.  Every thread of the block calls it.
__device__ __forceinline__ int rs_scan_digits(int v, int* wsum) {
  const int lane = threadIdx.x & (EB_WAVE - 1), wid = threadIdx.x / EB_WAVE;
  int x = v;
#pragma unroll
  for (int off = 1; off < EB_WAVE; off <<= 1) {
    const int y = __shfl_up(x, off, EB_WAVE);
    if (lane >= off) x += y;
  }
  if (lane == EB_WAVE - 1 && wid < RS_RADIX / EB_WAVE) wsum[wid] = x;
  __syncthreads();
  int before = 0;
  for (int w = 0; w < wid && w < RS_RADIX / EB_WAVE; ++w) before += wsum[w];
  __syncthreads();  // wsum is free again for the next call
  return (int)threadIdx.x < RS_RADIX ? before + x - v : 0;
}

// One tile per block.  Wave w ranks its own RS_SUB consecutive elements in RS_ROUNDS rounds
// (ballots find the equal digits of a round; per-wave digit counters in LDS carry the rounds,
// so the ranking needs no block barrier); the waves' counts are then summed per digit, the
// tile is reordered by digit in LDS, and the threads write it out in that order, so each
// digit's run of the tile goes to consecutive addresses.  Stable: within a digit the output
// follows the input order (tile, then wave, then round, then lane).
__global__ __launch_bounds__(RS_BT) void k_rs_scatter(const uint32_t* __restrict__ kin,
                                                      const int32_t* __restrict__ vin, uint32_t* __restrict__ kout,
                                                      int32_t* __restrict__ vout, int n, int shift, int nblk,
                                                      const int* __restrict__ pre, const int* __restrict__ tot) {
  __shared__ int wc[RS_BW][RS_RADIX];  // per-wave digit counts, then each wave's offset within the digit
  __shared__ int dglob[RS_RADIX];      // output position of the tile's first element of digit d
  __shared__ int dloc[RS_RADIX];       // tile-local position of digit d's first element
  __shared__ int wsum[RS_RADIX / EB_WAVE];
  __shared__ uint32_t sk[RS_TILE];
  __shared__ int32_t sv[RS_TILE];
  const int lane = threadIdx.x & (EB_WAVE - 1);
  const int wid = threadIdx.x / EB_WAVE;
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  const int tile = blockIdx.x * RS_TILE;
  const int w0 = tile + wid * RS_SUB;
  uint32_t kr[RS_ROUNDS];
  int32_t vr[RS_ROUNDS];
  int rk[RS_ROUNDS];
#pragma unroll
  for (int r = 0; r < RS_ROUNDS; ++r) {
    const int e = w0 + r * EB_WAVE + lane;
    kr[r] = e < n ? kin[e] : 0u;
    vr[r] = e < n ? vin[e] : 0;
  }
  for (int i = threadIdx.x; i < RS_BW * RS_RADIX; i += RS_BT) (&wc[0][0])[i] = 0;
  {
    // digits below d over all tiles, then this tile's place among d's tiles
    const int d = threadIdx.x;
    const int td = d < RS_RADIX ? tot[d] : 0;
    const int below = rs_scan_digits(td, wsum);  // also orders the clearing of wc before its use
    if (d < RS_RADIX) dglob[d] = below + pre[(size_t)d * nblk + blockIdx.x];
  }
#pragma unroll
  for (int r = 0; r < RS_ROUNDS; ++r) {
    const int e = w0 + r * EB_WAVE + lane;
    const bool valid = e < n;
    const int dig = valid ? (int)((kr[r] >> shift) & (RS_RADIX - 1)) : 0;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < RS_BITS; ++b) {
      const bool bit = (dig >> b) & 1;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const int rank = __popcll(peers & lt);
    // every lane reads the count before the digit's first lane raises it (one wave: the LDS
    // read instruction completes for all lanes before the write instruction)
    const int c = wc[wid][dig];
    rk[r] = c + rank;
    if (valid && rank == 0) wc[wid][dig] = c + __popcll(peers);
  }
  __syncthreads();
  int cnt = 0;
  if (threadIdx.x < RS_RADIX) {
    const int d = threadIdx.x;
    for (int w = 0; w < RS_BW; ++w) {
      const int c = wc[w][d];
      wc[w][d] = cnt;
      cnt += c;
    }
  }
  const int loc = rs_scan_digits(cnt, wsum);
  if (threadIdx.x < RS_RADIX) dloc[threadIdx.x] = loc;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RS_ROUNDS; ++r) {
    const int e = w0 + r * EB_WAVE + lane;
    if (e < n) {
      const int dig = (int)((kr[r] >> shift) & (RS_RADIX - 1));
      const int at = dloc[dig] + wc[wid][dig] + rk[r];
      sk[at] = kr[r];
      sv[at] = vr[r];
    }
  }
  __syncthreads();
  const int nt = min(RS_TILE, n - tile);
  for (int i = threadIdx.x; i < nt; i += RS_BT) {
    const uint32_t k = sk[i];
    const int dig = (int)((k >> shift) & (RS_RADIX - 1));
    const int at = dglob[dig] + (i - dloc[dig]);
    kout[at] = k;
    vout[at] = sv[i];
  }
}

// Exclusive scan of n ints in one block of 1024 threads; out[n] receives the total when
// out has n+1 entries (write_total).
__global__ __launch_bounds__(1024) void k_scan1(const int* __restrict__ in, int* __restrict__ out, int n,
                                               int write_total) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int chunk = (n + 1023) / 1024;
  const int s = t * chunk;
  const int e = min(n, s + chunk);
  int sum = 0;
  for (int i = s; i < e; ++i) sum += in[i];
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = (t >= off) ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - sum;
  for (int i = s; i < e; ++i) {
    int x = in[i];
    out[i] = run;
    run += x;
  }
  if (write_total && t == 1023) out[n] = part[1023];
}

// Three-phase exclusive scan for larger arrays: per-tile sums, a scan of the tile sums, then
// each tile scanned again with its offset.  Tiles of SC_TILE ints are staged through LDS so
// every global access is coalesced.
constexpr int SC_THREADS = 256;
constexpr int SC_ITEMS = 8;
constexpr int SC_TILE = SC_THREADS * SC_ITEMS;

__global__ __launch_bounds__(SC_THREADS) void k_scan_tilesum(const int* __restrict__ in, int n, int* __restrict__ sums) {
  __shared__ int red[SC_THREADS];
  const int base = blockIdx.x * SC_TILE;
  int s = 0;
  for (int i = threadIdx.x; i < SC_TILE; i += SC_THREADS) {
    const int e = base + i;
    if (e < n) s += in[e];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = SC_THREADS / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(SC_THREADS) void k_scan_tile(const int* __restrict__ in, int* __restrict__ out, int n,
                                                          const int* __restrict__ offs, int write_total) {
  __shared__ int tile[SC_TILE];
  __shared__ int part[SC_THREADS];
  const int base = blockIdx.x * SC_TILE;
  for (int i = threadIdx.x; i < SC_TILE; i += SC_THREADS) {
    const int e = base + i;
    tile[i] = e < n ? in[e] : 0;
  }
  __syncthreads();
  int loc[SC_ITEMS];
  int s = 0;
#pragma unroll
  for (int j = 0; j < SC_ITEMS; ++j) {
    loc[j] = s;
    s += tile[threadIdx.x * SC_ITEMS + j];
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < SC_THREADS; off <<= 1) {
    const int v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  const int run = offs[blockIdx.x] + part[threadIdx.x] - s;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < SC_ITEMS; ++j) tile[threadIdx.x * SC_ITEMS + j] = run + loc[j];
  __syncthreads();
  for (int i = threadIdx.x; i < SC_TILE; i += SC_THREADS) {
    const int e = base + i;
    if (e < n) out[e] = tile[i];
  }
  if (write_total && blockIdx.x == gridDim.x - 1 && threadIdx.x == SC_THREADS - 1)
    out[n] = offs[blockIdx.x] + part[SC_THREADS - 1];
}

size_t scan_temp_ints(int n) { return (size_t)2 * ((n + SC_TILE - 1) / SC_TILE) + 1; }

// Exclusive scan; `tmp` holds scan_temp_ints(n) ints owned by the caller.
int scan_exclusive(const int* in, int* out, int n, bool write_total, int* tmp, hipStream_t st) {
  const int nt = (n + SC_TILE - 1) / SC_TILE;
  if (nt <= 1 || tmp == nullptr) {
    hipLaunchKernelGGL(k_scan1, dim3(1), dim3(1024), 0, st, in, out, n, write_total ? 1 : 0);
    EB_HIP_CHECK(hipGetLastError());
    return EB_OK;
  }
  int* sums = tmp;
  int* offs = tmp + nt;
  hipLaunchKernelGGL(k_scan_tilesum, dim3(nt), dim3(SC_THREADS), 0, st, in, n, sums);
  hipLaunchKernelGGL(k_scan1, dim3(1), dim3(1024), 0, st, sums, offs, nt, 0);
  hipLaunchKernelGGL(k_scan_tile, dim3(nt), dim3(SC_THREADS), 0, st, in, out, n, offs, write_total ? 1 : 0);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

size_t radix_sort_temp_ints(int n) {
  int nblk = (n + RS_TILE - 1) / RS_TILE;
  return (size_t)2 * RS_RADIX * nblk + RS_RADIX;
}

// Sorts (k0, v0) by the low `nbits` of the keys; the result ends in (k0, v0).  k1/v1 are
// scratch of n entries, tmp holds radix_sort_temp_ints(n) ints.
int radix_sort_pairs(uint32_t* k0, int32_t* v0, uint32_t* k1, int32_t* v1, int n, int nbits, int* tmp,
                     hipStream_t st, bool hist0_ready) {
  if (n <= 0) return EB_OK;
  const int nblk = (n + RS_TILE - 1) / RS_TILE;
  int* hist = tmp;
  int* pre = tmp + RS_RADIX * nblk;
  int* tot = pre + RS_RADIX * nblk;
  int passes = (nbits + RS_BITS - 1) / RS_BITS;
  if (passes == 0) passes = 1;
  if (passes & 1) passes += 1;  // an even number of passes leaves the result in (k0, v0)
  uint32_t* ka = k0;
  int32_t* va = v0;
  uint32_t* kb = k1;
  int32_t* vb = v1;
  for (int p = 0; p < passes; ++p) {
    const int shift = p * RS_BITS;
    if (shift < 32) {
      if (p > 0 || !hist0_ready)
        hipLaunchKernelGGL(k_rs_hist, dim3(nblk), dim3(RS_BT), 0, st, ka, n, shift, nblk, hist);
      hipLaunchKernelGGL(k_rs_prefix, dim3(RS_RADIX), dim3(RS_THREADS), 0, st, hist, nblk, pre, tot);
      hipLaunchKernelGGL(k_rs_scatter, dim3(nblk), dim3(RS_BT), 0, st, ka, va, kb, vb, n, shift, nblk, pre, tot);
      EB_HIP_CHECK(hipGetLastError());
    } else {
      EB_HIP_CHECK(hipMemcpyAsync(kb, ka, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, st));
      EB_HIP_CHECK(hipMemcpyAsync(vb, va, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, st));
    }
    uint32_t* tk = ka; ka = kb; kb = tk;
    int32_t* tv = va; va = vb; vb = tv;
  }
  return EB_OK;
}

}  // namespace eb
