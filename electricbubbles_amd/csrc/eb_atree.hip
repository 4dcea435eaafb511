// eb_atree.hip -- DAPIF2's adaptive quadtree on the GPU (include/eb_atree.h).
//
// Every step of the reference's serial construction is restated as data-parallel passes
// that produce the same integers:
//   * DAIBNB's per-particle child choice is a comparison with the parent centre; the counts
//     are integer atomics (order-free).
//   * DAFAMI numbers the children of level-L boxes in the order of their parents: an
//     exclusive scan of the "splits" flags.
//   * DACOMP's SHIFT table is an exclusive scan of the "empty" flags.
//   * DAIJAT's IAT chain is an exclusive scan of NBAT+1 over childless boxes; the particle
//     order inside a leaf is the index order, i.e. a stable sort by box.
//   * DALI12 runs level by level (a box needs its parent's colleagues): count, scan, write.
//   * DALI34's per-box breadth-first scan runs for all boxes at once, one depth per pass,
//     over a global frontier kept in (box, BFS) order; list entries of each depth are
//     appended, and a stable sort by box at the end restores the reference's per-box
//     concatenation (depth after depth, BFS order within a depth).
// The only floating-point values are the scaled coordinates (one subtraction and one
// product per point, as in dapif1.f:765-767) and dyadic box centres and distances.
#include <algorithm>
#include <vector>

#include "../../include/eb_atree.h"
#include "eb_internal.h"

namespace {

constexpr int TPB = 256;
constexpr int AT_MAXLEV = 120;
inline int nblk(long n) { return (int)std::max<long>(1, (n + TPB - 1) / TPB); }

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t cap = 0;
};

// Grows b to hold `need` elements, keeping the first `keep`, zeroing the rest.
template <typename T>
int grow(DBuf<T>& b, size_t need, size_t keep, hipStream_t st) {
  if (need <= b.cap) return EB_OK;
  size_t c = b.cap ? b.cap : 256;
  while (c < need) c *= 2;
  T* q = nullptr;
  if (hipMalloc((void**)&q, sizeof(T) * c) != hipSuccess) {
    eb_set_error("adaptive tree: hipMalloc(%zu bytes) failed", sizeof(T) * c);
    return EB_ERR_NOMEM;
  }
  EB_HIP_CHECK(hipMemsetAsync(q, 0, sizeof(T) * c, st));
  if (b.p) {
    if (keep) EB_HIP_CHECK(hipMemcpyAsync(q, b.p, sizeof(T) * keep, hipMemcpyDeviceToDevice, st));
    EB_HIP_CHECK(hipStreamSynchronize(st));
    EB_HIP_CHECK(hipFree(b.p));
  }
  b.p = q;
  b.cap = c;
  return EB_OK;
}

template <typename T>
void release(DBuf<T>& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
}

// ------------------------------------------------------------------------------------------
// scaling (dapif1.f:745-770)
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_at_minmax(const double* __restrict__ x, const double* __restrict__ y,
                                                   int n, double4* __restrict__ part) {
  __shared__ double4 s[TPB];
  double4 a = make_double4(1.0 / 0.0, -1.0 / 0.0, 1.0 / 0.0, -1.0 / 0.0);
  for (int i = blockIdx.x * TPB + threadIdx.x; i < n; i += gridDim.x * TPB) {
    a.x = fmin(a.x, x[i]);
    a.y = fmax(a.y, x[i]);
    a.z = fmin(a.z, y[i]);
    a.w = fmax(a.w, y[i]);
  }
  s[threadIdx.x] = a;
  __syncthreads();
  for (int off = TPB / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      const double4 b = s[threadIdx.x + off];
      a = s[threadIdx.x];
      s[threadIdx.x] = make_double4(fmin(a.x, b.x), fmax(a.y, b.y), fmin(a.z, b.z), fmax(a.w, b.w));
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// RMINX .. RMAXY start at +-1.0E30 (a REAL*4 constant); RSCAL = BOXSIZ / max range * 0.99
// with the REAL*4 0.99 widened to double; XONEW, YONEW are the box centre.
__global__ __launch_bounds__(TPB) void k_at_geom(const double4* __restrict__ part, int np, double* __restrict__ geom) {
#pragma clang fp contract(off)
  __shared__ double4 s[TPB];
  const double big = (double)1.0e30f;
  double4 a = make_double4(big, -big, big, -big);
  for (int i = threadIdx.x; i < np; i += TPB) {
    const double4 b = part[i];
    a = make_double4(fmin(a.x, b.x), fmax(a.y, b.y), fmin(a.z, b.z), fmax(a.w, b.w));
  }
  s[threadIdx.x] = a;
  __syncthreads();
  for (int off = TPB / 2; off > 0; off >>= 1) {  // min / max are exact: any order gives the same
    if (threadIdx.x < off) {
      const double4 b = s[threadIdx.x + off];
      a = s[threadIdx.x];
      s[threadIdx.x] = make_double4(fmin(a.x, b.x), fmax(a.y, b.y), fmin(a.z, b.z), fmax(a.w, b.w));
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double mnx = s[0].x, mxx = s[0].y, mny = s[0].z, mxy = s[0].w;
  const double rsx = mxx - mnx, rsy = mxy - mny;
  const double xo = (mxx + mnx) / 2.0, yo = (mxy + mny) / 2.0;
  const double rs = rsx >= rsy ? rsx : rsy;
  geom[0] = 128.0 / rs * (double)0.99f;
  geom[1] = xo;
  geom[2] = yo;
}

__global__ void k_at_scale(const double* __restrict__ x, const double* __restrict__ y, int n,
                           const double* __restrict__ geom, double* __restrict__ xs, double* __restrict__ ys) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r = geom[0];
  xs[i] = (x[i] - geom[1]) * r;
  ys[i] = (y[i] - geom[2]) * r;
}

// ------------------------------------------------------------------------------------------
// DAASSI loop: DAINIA (dapif1.f:1262-1300), DAIBNB (1303-1370), DAFAMI (1373-1453)
// Arrays are 0-based by box (box b stored at b-1) and hold the reference's 1-based values.
// ------------------------------------------------------------------------------------------
__global__ void k_at_init(double* __restrict__ xb, double* __restrict__ yb) {
  const int b = threadIdx.x;
  if (b >= 4) return;
  const double d = 0.25 * 128.0;
  xb[b] = (b & 1) ? d : -d;   // NW, NE, SW, SE
  yb[b] = (b & 2) ? -d : d;
}

__global__ void k_at_zero(int* __restrict__ a, int lo, int hi) {  // a[lo..hi] = 0
  const int i = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= hi) a[i] = 0;
}

// Child quadrant: x < xp is west, y < yp is south; NW = nc, NE = nc+1, SW = nc+2, SE = nc+3.
// The counts are aggregated over the wave first: at the coarse levels most points of a wave
// fall into the same few boxes, and one atomic per distinct box replaces one per point.
__global__ void k_at_assign(int n, const double* __restrict__ xs, const double* __restrict__ ys,
                            int* __restrict__ ibox, const double* __restrict__ xb, const double* __restrict__ yb,
                            int* __restrict__ nbat, const int* __restrict__ nchi, int napb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int key = -1;  // child box (1-based) this point moves to, or -1
  if (i < n) {
    const int ib = ibox[i];
    double xp = 0.0, yp = 0.0;
    int nc = 1;
    bool move = true;
    if (ib != 0) {
      move = nbat[ib - 1] > napb;
      if (move) {
        xp = xb[ib - 1];
        yp = yb[ib - 1];
        nc = nchi[ib - 1];
      }
    }
    if (move) {
      const double x = xs[i], y = ys[i];
      const int q = (x < xp) ? ((y < yp) ? 2 : 0) : ((y < yp) ? 3 : 1);
      key = nc + q;
      ibox[i] = key;
    }
  }
  const int lane = threadIdx.x & 63;
  unsigned long long todo = __ballot(key >= 0);
  for (int round = 0; round < 8 && todo; ++round) {
    const int leader = __ffsll((long long)todo) - 1;
    const int lk = __shfl(key, leader);
    const unsigned long long same = __ballot(key == lk) & todo;
    if (lane == leader) atomicAdd(&nbat[lk - 1], __popcll(same));
    todo &= ~same;
  }
  if (key >= 0 && ((todo >> lane) & 1ull)) atomicAdd(&nbat[key - 1], 1);  // many distinct boxes
}

__global__ void k_at_split_flag(const int* __restrict__ nbat, int lo, int m, int napb, int* __restrict__ flag) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) flag[j] = nbat[lo + j] > napb ? 1 : 0;
}

__global__ void k_at_children(int lo, int m, int nbox, const int* __restrict__ flag, const int* __restrict__ pos,
                              int* __restrict__ npar, int* __restrict__ nchi, double* __restrict__ xb,
                              double* __restrict__ yb, double d) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const int b = lo + j;
  if (!flag[j]) {
    nchi[b] = 0;
    return;
  }
  const int c = nbox + 4 * pos[j];
  nchi[b] = c + 1;
  const double x = xb[b], y = yb[b];
  for (int k = 0; k < 4; ++k) {
    npar[c + k] = b + 1;
    xb[c + k] = (k & 1) ? x + d : x - d;
    yb[c + k] = (k & 2) ? y - d : y + d;
  }
}

// ------------------------------------------------------------------------------------------
// DACOMP (dapif1.f:1456-1539).  shift[b] = SHIFT(b+1) = number of empty boxes before box
// b+1; shift[nbox] = SHIFT(NBOX+1).  SHIFT(0), read for NCHI = 0 and NPAR = 0, is 0.
// ------------------------------------------------------------------------------------------
__global__ void k_at_empty(const int* __restrict__ nbat, int nbox, int* __restrict__ e) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nbox) e[b] = nbat[b] == 0 ? 1 : 0;
}

__device__ __forceinline__ int shift_of(const int* __restrict__ shift, int v) { return v == 0 ? 0 : shift[v - 1]; }

__global__ void k_at_compact(int nbox, const int* __restrict__ shift, const double* __restrict__ xb,
                             const double* __restrict__ yb, const int* __restrict__ nbat,
                             const int* __restrict__ npar, const int* __restrict__ nchi, double* __restrict__ xb2,
                             double* __restrict__ yb2, int* __restrict__ nbat2, int* __restrict__ npar2,
                             int* __restrict__ nchi2) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbox || nbat[b] <= 0) return;
  const int nb = b - shift[b];
  xb2[nb] = xb[b];
  yb2[nb] = yb[b];
  nbat2[nb] = nbat[b];
  npar2[nb] = npar[b] - shift_of(shift, npar[b]);
  nchi2[nb] = nchi[b] - shift_of(shift, nchi[b]);
}

__global__ void k_at_ibox_shift(int n, const int* __restrict__ shift, int* __restrict__ ibox) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ibox[i] -= shift_of(shift, ibox[i]);
}

__global__ void k_at_nl_shift(int cnt, const int* __restrict__ shift, int* __restrict__ nl) {
  const int k = threadIdx.x;
  if (k < cnt) nl[k] -= shift_of(shift, nl[k]);
}

// ------------------------------------------------------------------------------------------
// DAINDB (dapif1.f:1542-1606) and DAIJAT (dapif1.f:1609-1680)
// ------------------------------------------------------------------------------------------
__global__ void k_at_indb(int nbox, int nlev, const int* __restrict__ nl, const int* __restrict__ nbat,
                          int* __restrict__ nchi, int napb, int* __restrict__ indb, int* __restrict__ maxp) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbox) return;
  int lev = 1;
  while (lev < nlev && nl[lev] <= b + 1) ++lev;  // nl[k] = NL(k+1)
  const int nb = nbat[b];
  if (lev == nlev) {
    indb[b] = 1;
    nchi[b] = 0;
    atomicMax(maxp, nb);
  } else if (nb <= napb) {
    indb[b] = 1;
    atomicMax(maxp, nb);
  } else {
    indb[b] = 0;
  }
}

__global__ void k_at_iat_weight(int nbox, const int* __restrict__ indb, const int* __restrict__ nbat,
                                int* __restrict__ w) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nbox) w[b] = indb[b] ? nbat[b] + 1 : 0;
}

// IAT(b) = 1 + sum over earlier childless boxes of (NBAT + 1); JAT(IAT(b)) = NBAT(b).
__global__ void k_at_iat(int nbox, const int* __restrict__ indb, const int* __restrict__ nbat,
                         const int* __restrict__ ex, int* __restrict__ iat, int* __restrict__ jat) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbox) return;
  if (indb[b]) {
    iat[b] = 1 + ex[b];
    jat[ex[b]] = nbat[b];
  } else {
    iat[b] = -nbat[b];
  }
}

__global__ void k_at_keys(int n, const int* __restrict__ ibox, uint32_t* __restrict__ k, int32_t* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  k[i] = (uint32_t)ibox[i];
  v[i] = i;
}

__global__ void k_at_first(int n, const uint32_t* __restrict__ k, int* __restrict__ first) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  if (s == 0 || k[s] != k[s - 1]) first[k[s] - 1] = s;
}

__global__ void k_at_jat(int n, const uint32_t* __restrict__ k, const int32_t* __restrict__ v,
                         const int* __restrict__ first, const int* __restrict__ iat, int* __restrict__ jat,
                         int* __restrict__ err) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int b = (int)k[s] - 1;
  if (iat[b] <= 0) {  // a point outside every childless box: the tree is inconsistent
    atomicOr(err, 2);
    return;
  }
  jat[iat[b] + (s - first[b])] = v[s] + 1;
}

// ------------------------------------------------------------------------------------------
// DALI12 + DAIN12 (dapif2.f:3-210)
// l1[k] = L1(k+1), l2[k] = L2(k+1), jl[k] = LL(k+1).
// ------------------------------------------------------------------------------------------
__global__ void k_at_in12(int imin, int imax, int* __restrict__ l1, int* __restrict__ l2, int* __restrict__ jl) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  l2[0] = 1;
  l1[0] = 1;
  for (int i = imin; i <= imax; ++i) {
    int lc = l2[i - 1];
    for (int j = imin; j <= imax; ++j) {
      if (i == j) continue;
      jl[lc - 1] = j;
      ++lc;
    }
    l2[i] = lc;
    l1[i] = lc;
  }
}

constexpr int LI_MAX1 = 27, LI_MAX2 = 8;

// Interaction list (inli) and colleagues (col) of box I from its parent's colleagues.
__device__ bool li12_box(int I, double d, const double* __restrict__ xb, const double* __restrict__ yb,
                         const int* __restrict__ npar, const int* __restrict__ nchi, const int* __restrict__ l1,
                         const int* __restrict__ l2, const int* __restrict__ jl, int* inli, int& nli, int* col,
                         int& nco) {
  const int np = npar[I - 1];
  const int lpmin = l2[np - 1], lpmax = l1[np] - 1;
  const double x = xb[I - 1], y = yb[I - 1];
  nli = 0;
  nco = 0;
  bool ok = true;
  for (int jco = lpmin; jco <= lpmax; ++jco) {
    const int ico = jl[jco - 1];
    int icok = nchi[ico - 1];
    if (icok == 0) continue;
    while (npar[icok - 1] == ico) {
      if (fabs(x - xb[icok - 1]) <= d && fabs(y - yb[icok - 1]) <= d) {
        if (nco < LI_MAX2) col[nco++] = icok; else ok = false;
      } else {
        if (nli < LI_MAX1) inli[nli++] = icok; else ok = false;
      }
      ++icok;
    }
  }
  for (int nbro = nchi[np - 1]; npar[nbro - 1] == np; ++nbro) {
    if (nbro == I) continue;
    if (nco < LI_MAX2) col[nco++] = nbro; else ok = false;
  }
  return ok;
}

__global__ void k_at_li12_count(int lo, int m, double d, const double* __restrict__ xb,
                                const double* __restrict__ yb, const int* __restrict__ npar,
                                const int* __restrict__ nchi, const int* __restrict__ l1, const int* __restrict__ l2,
                                const int* __restrict__ jl, int* __restrict__ cnt, int* __restrict__ err) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  int inli[LI_MAX1], col[LI_MAX2], nli, nco;
  if (!li12_box(lo + j, d, xb, yb, npar, nchi, l1, l2, jl, inli, nli, col, nco)) atomicOr(err, 1);
  cnt[j] = nli + nco;
}

__global__ void k_at_li12_write(int lo, int m, double d, const double* __restrict__ xb,
                                const double* __restrict__ yb, const int* __restrict__ npar,
                                const int* __restrict__ nchi, int* __restrict__ l1, int* __restrict__ l2,
                                int* __restrict__ jl, const int* __restrict__ ex) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const int I = lo + j;
  int inli[LI_MAX1], col[LI_MAX2], nli, nco;
  li12_box(I, d, xb, yb, npar, nchi, l1, l2, jl, inli, nli, col, nco);
  const int lr = l1[lo - 1] + ex[j];  // L1(lo) from the previous level, + this box's offset
  for (int k = 0; k < nli; ++k) jl[lr - 1 + k] = inli[k];
  l2[I - 1] = lr + nli;
  for (int k = 0; k < nco; ++k) jl[lr + nli - 1 + k] = col[k];
  l1[I] = lr + nli + nco;
}

// ------------------------------------------------------------------------------------------
// DALI34 + DASCAN (dapif2.f:212-474)
// ------------------------------------------------------------------------------------------
__global__ void k_at_li34_init(int nbox, const int* __restrict__ indb, const int* __restrict__ l1,
                               const int* __restrict__ l2, const int* __restrict__ jl, const double* __restrict__ xb,
                               const double* __restrict__ yb, int* __restrict__ cnt, double* __restrict__ d1) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbox) return;
  const int l2i = l2[b], l1i = l1[b + 1] - 1;
  cnt[b] = 0;
  if (l1i < l2i || indb[b] == 0) return;
  bool nnbar = false;
  for (int j = l2i; j <= l1i; ++j) nnbar |= indb[jl[j - 1] - 1] == 0;
  if (!nnbar) return;
  cnt[b] = l1i - l2i + 1;
  const int c = jl[l2i - 1];
  const double xd = fabs(xb[b] - xb[c - 1]), yd = fabs(yb[b] - yb[c - 1]);
  d1[b] = xd >= yd ? xd : yd;
}

__global__ void k_at_li34_front(int nbox, const int* __restrict__ cnt, const int* __restrict__ ex,
                                const int* __restrict__ l2, const int* __restrict__ jl, int* __restrict__ fo,
                                int* __restrict__ fn) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbox) return;
  for (int k = 0; k < cnt[b]; ++k) {
    fo[ex[b] + k] = b + 1;
    fn[ex[b] + k] = jl[l2[b] - 1 + k];
  }
}

// One DASCAN depth for every frontier entry (owner box, descendant): list 4 if not adjacent,
// list 3 if adjacent and childless, else its children become the next frontier.
__global__ void k_at_li34_classify(int nf, const int* __restrict__ fo, const int* __restrict__ fn,
                                   const double* __restrict__ xb, const double* __restrict__ yb,
                                   const int* __restrict__ indb, const int* __restrict__ nchi,
                                   const int* __restrict__ npar, const double* __restrict__ d1, double fac,
                                   int* __restrict__ f3, int* __restrict__ f4, int* __restrict__ fc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nf) return;
  const int I = fo[p], ibo = fn[p];
  const double d = d1[I - 1] * fac;
  const double xd = fabs(xb[I - 1] - xb[ibo - 1]), yd = fabs(yb[I - 1] - yb[ibo - 1]);
  int a3 = 0, a4 = 0, c = 0;
  if (xd > d || yd > d) {
    a4 = 1;
  } else if (indb[ibo - 1] == 1) {
    a3 = 1;
  } else {
    for (int k = nchi[ibo - 1]; npar[k - 1] == ibo; ++k) ++c;
  }
  f3[p] = a3;
  f4[p] = a4;
  fc[p] = c;
}

__global__ void k_at_li34_emit(int nf, const int* __restrict__ fo, const int* __restrict__ fn,
                               const int* __restrict__ f3, const int* __restrict__ f4, const int* __restrict__ fc,
                               const int* __restrict__ e3, const int* __restrict__ e4, const int* __restrict__ ec,
                               int base3, int base4, int* __restrict__ o3o, int* __restrict__ o3n,
                               int* __restrict__ o4o, int* __restrict__ o4n, int* __restrict__ nfo,
                               int* __restrict__ nfn, const int* __restrict__ nchi) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nf) return;
  const int I = fo[p], ibo = fn[p];
  if (f3[p]) {
    o3o[base3 + e3[p]] = I;
    o3n[base3 + e3[p]] = ibo;
  }
  if (f4[p]) {
    o4o[base4 + e4[p]] = I;
    o4n[base4 + e4[p]] = ibo;
  }
  const int c = fc[p];
  if (c) {
    const int k0 = nchi[ibo - 1];
    for (int k = 0; k < c; ++k) {
      nfo[ec[p] + k] = I;
      nfn[ec[p] + k] = k0 + k;
    }
  }
}

__global__ void k_at_hist(int n, const int* __restrict__ owner, int* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&cnt[owner[i] - 1], 1);
}

__global__ void k_at_ptr(int nb1, const int* __restrict__ ex, int* __restrict__ ptr) {  // L(b) = 1 + ex[b]
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb1) ptr[b] = 1 + ex[b];
}

int bits_for(long v) {
  int b = 1;
  while ((1L << b) <= v) ++b;
  return b;
}

}  // namespace

struct eb_atree_s {
  int nat = 0, napb = 0, nbox = 0, nlev = 0, len = 0, len3 = 0, len4 = 0, maxp = 0, jatlen = 0;
  hipStream_t st = nullptr;
  DBuf<double> xs, ys, geom, xb, yb;
  DBuf<int> nl, ibox, nbat, npar, nchi, indb, iat, jat, l1, l2, jl, l3, l4, jl3, jl4;
  // scratch
  DBuf<int> s_a, s_b, s_c, s_d, s_e, s_f, s_g, s_h, s_tmp;
  DBuf<double> s_d1;
  DBuf<double4> s_part;
  DBuf<uint32_t> s_k0, s_k1;
  DBuf<int32_t> s_v0, s_v1;
  void free_scratch() {
    release(s_a); release(s_b); release(s_c); release(s_d); release(s_e); release(s_f); release(s_g);
    release(s_h); release(s_tmp); release(s_d1); release(s_part); release(s_k0); release(s_k1);
    release(s_v0); release(s_v1);
  }
  ~eb_atree_s() {
    free_scratch();
    release(xs); release(ys); release(geom); release(xb); release(yb); release(nl); release(ibox);
    release(nbat); release(npar); release(nchi); release(indb); release(iat); release(jat); release(l1);
    release(l2); release(jl); release(l3); release(l4); release(jl3); release(jl4);
  }
};

namespace {

#define AT_TRY(expr)          \
  do {                        \
    int _rc = (expr);         \
    if (_rc != EB_OK) return _rc; \
  } while (0)

int scan(eb_atree_s* T, const int* in, int* out, int n, bool total) {
  AT_TRY(grow(T->s_tmp, eb::scan_temp_ints(n) + 1, 0, T->st));
  return eb::scan_exclusive(in, out, n, total, T->s_tmp.p, T->st);
}

int read_int(eb_atree_s* T, const int* dptr, int* out) {
  EB_HIP_CHECK(hipMemcpyAsync(out, dptr, sizeof(int), hipMemcpyDeviceToHost, T->st));
  EB_HIP_CHECK(hipStreamSynchronize(T->st));
  return EB_OK;
}

// Stable sort of (owner, node) pairs by owner; fills ptr (nbox + 1 pointers, 1-based) and
// list (the nodes, grouped by owner).
int group_by_owner(eb_atree_s* T, int n, const int* owner, const int* node, DBuf<int>& ptr, DBuf<int>& list) {
  const int nbox = T->nbox;
  hipStream_t st = T->st;
  AT_TRY(grow(ptr, nbox + 2, 0, st));
  AT_TRY(grow(list, std::max(n, 1), 0, st));
  AT_TRY(grow(T->s_g, nbox + 2, 0, st));
  AT_TRY(grow(T->s_h, nbox + 2, 0, st));
  EB_HIP_CHECK(hipMemsetAsync(T->s_g.p, 0, sizeof(int) * (nbox + 1), st));
  if (n > 0) hipLaunchKernelGGL(k_at_hist, dim3(nblk(n)), dim3(TPB), 0, st, n, owner, T->s_g.p);
  AT_TRY(scan(T, T->s_g.p, T->s_h.p, nbox, true));
  hipLaunchKernelGGL(k_at_ptr, dim3(nblk(nbox + 1)), dim3(TPB), 0, st, nbox + 1, T->s_h.p, ptr.p);
  if (n > 0) {
    AT_TRY(grow(T->s_k0, n, 0, st));
    AT_TRY(grow(T->s_k1, n, 0, st));
    AT_TRY(grow(T->s_v0, n, 0, st));
    AT_TRY(grow(T->s_v1, n, 0, st));
    EB_HIP_CHECK(hipMemcpyAsync(T->s_k0.p, owner, sizeof(int) * n, hipMemcpyDeviceToDevice, st));
    EB_HIP_CHECK(hipMemcpyAsync(T->s_v0.p, node, sizeof(int) * n, hipMemcpyDeviceToDevice, st));
    AT_TRY(grow(T->s_tmp, eb::radix_sort_temp_ints(n) + 1, 0, st));
    AT_TRY(eb::radix_sort_pairs(T->s_k0.p, T->s_v0.p, T->s_k1.p, T->s_v1.p, n, bits_for(nbox + 1), T->s_tmp.p,
                                st));
    EB_HIP_CHECK(hipMemcpyAsync(list.p, T->s_v0.p, sizeof(int) * n, hipMemcpyDeviceToDevice, st));
  }
  return EB_OK;
}

int build(eb_atree_s* T, const double* x, const double* y, int lmx, double eps7, long long nboxmx, int32_t* ierr) {
  const int n = T->nat, napb = T->napb;
  hipStream_t st = T->st;
  // ---- scaling
  AT_TRY(grow(T->xs, n, 0, st));
  AT_TRY(grow(T->ys, n, 0, st));
  AT_TRY(grow(T->geom, 4, 0, st));
  const int np = std::min(1024, nblk(n));
  AT_TRY(grow(T->s_part, np, 0, st));
  hipLaunchKernelGGL(k_at_minmax, dim3(np), dim3(TPB), 0, st, x, y, n, T->s_part.p);
  hipLaunchKernelGGL(k_at_geom, dim3(1), dim3(TPB), 0, st, T->s_part.p, np, T->geom.p);
  hipLaunchKernelGGL(k_at_scale, dim3(nblk(n)), dim3(TPB), 0, st, x, y, n, T->geom.p, T->xs.p, T->ys.p);
  EB_HIP_CHECK(hipGetLastError());

  // ---- DAASSI
  AT_TRY(grow(T->ibox, n, 0, st));
  DBuf<double> xb, yb;
  DBuf<int> nbat, npar, nchi;
  struct Guard {  // level buffers, replaced by the compacted tables
    hipStream_t st;
    DBuf<double>*a, *b;
    DBuf<int>*c, *d, *e;
    ~Guard() {
      (void)hipStreamSynchronize(st);
      release(*a); release(*b); release(*c); release(*d); release(*e);
    }
  } guard{st, &xb, &yb, &nbat, &npar, &nchi};
  size_t cap = std::max<size_t>(1024, (size_t)(4.0 * n / (napb + 1)) + 64);
  AT_TRY(grow(xb, cap, 0, st));
  AT_TRY(grow(yb, cap, 0, st));
  AT_TRY(grow(nbat, cap, 0, st));
  AT_TRY(grow(npar, cap, 0, st));
  AT_TRY(grow(nchi, cap, 0, st));
  hipLaunchKernelGGL(k_at_init, dim3(1), dim3(64), 0, st, xb.p, yb.p);
  std::vector<int> nl(AT_MAXLEV + 4, 0);  // nl[k] = NL(k), 1-based level index
  int nbox = 4, L = 1;
  nl[1] = 1;
  nl[2] = 5;
  bool splits = true;  // MAX = NAT > NAPB on entry
  for (;;) {
    if (L > lmx) {
      ierr[0] = EB_ATREE_TOO_DEEP;
      ierr[1] = L - 1;
      ierr[2] = lmx - 1;
      return EB_ATREE_TOO_DEEP;
    }
    if (ldexp(1.0, -(L - 1)) <= eps7) {
      ierr[0] = EB_ATREE_EPS_STOP;
      return EB_ATREE_EPS_STOP;
    }
    if (!splits) break;
    // DAIBNB
    const int zlo = nl[L] - 1, zhi = nl[L + 1] - 1;
    hipLaunchKernelGGL(k_at_zero, dim3(nblk(zhi - zlo + 1)), dim3(TPB), 0, st, nbat.p, zlo, zhi);
    hipLaunchKernelGGL(k_at_assign, dim3(nblk(n)), dim3(TPB), 0, st, n, T->xs.p, T->ys.p, T->ibox.p, xb.p, yb.p,
                       nbat.p, nchi.p, napb);
    // DAFAMI
    const int lo = nl[L] - 1, m = nbox - lo;
    AT_TRY(grow(T->s_a, m + 1, 0, st));
    AT_TRY(grow(T->s_b, m + 1, 0, st));
    hipLaunchKernelGGL(k_at_split_flag, dim3(nblk(m)), dim3(TPB), 0, st, nbat.p, lo, m, napb, T->s_a.p);
    AT_TRY(scan(T, T->s_a.p, T->s_b.p, m, true));
    int nsplit = 0;
    AT_TRY(read_int(T, T->s_b.p + m, &nsplit));
    if ((long long)nbox + 4LL * nsplit > nboxmx) {
      // DAFAMI adds four children per split box and stops at the first count past NBOXMX
      // (dapif1.f:1436-1441), with IERR(2) = (NBOXMX - NBOX) * 36 + NSP, NSP being an
      // undeclared local of DAFAMI (zero in the static storage the reference needs)
      const long long over = (long long)nbox + 4 * ((nboxmx - nbox) / 4 + 1);
      ierr[0] = EB_ATREE_WORKSPACE;
      ierr[1] = (int32_t)((nboxmx - over) * 36);
      return EB_ATREE_WORKSPACE;
    }
    const size_t need = (size_t)nbox + 4 * (size_t)nsplit + 2;
    AT_TRY(grow(xb, need, nbox, st));
    AT_TRY(grow(yb, need, nbox, st));
    AT_TRY(grow(nbat, need, nbox, st));
    AT_TRY(grow(npar, need, nbox, st));
    AT_TRY(grow(nchi, need, nbox, st));
    const double d = 0.25 / ldexp(1.0, L) * 128.0;
    hipLaunchKernelGGL(k_at_children, dim3(nblk(m)), dim3(TPB), 0, st, lo, m, nbox, T->s_a.p, T->s_b.p, npar.p,
                       nchi.p, xb.p, yb.p, d);
    EB_HIP_CHECK(hipGetLastError());
    nbox += 4 * nsplit;
    L += 1;
    if (L + 1 >= AT_MAXLEV + 3) {
      eb_set_error("adaptive tree: more than %d levels", AT_MAXLEV);
      return EB_ERR_ARG;
    }
    nl[L + 1] = nbox + 1;
    splits = nsplit > 0;
  }
  const int nlev = L - 1;

  // ---- DACOMP
  AT_TRY(grow(T->s_c, nbox + 1, 0, st));
  AT_TRY(grow(T->s_d, nbox + 2, 0, st));
  hipLaunchKernelGGL(k_at_empty, dim3(nblk(nbox)), dim3(TPB), 0, st, nbat.p, nbox, T->s_c.p);
  AT_TRY(scan(T, T->s_c.p, T->s_d.p, nbox, true));
  int nz = 0;
  AT_TRY(read_int(T, T->s_d.p + nbox, &nz));
  const int nb = nbox - nz;
  AT_TRY(grow(T->xb, nb + 1, 0, st));
  AT_TRY(grow(T->yb, nb + 1, 0, st));
  AT_TRY(grow(T->nbat, nb + 1, 0, st));
  AT_TRY(grow(T->npar, nb + 2, 0, st));  // NPAR(NBOX+1) = 0 is the end marker of child runs
  AT_TRY(grow(T->nchi, nb + 1, 0, st));
  hipLaunchKernelGGL(k_at_compact, dim3(nblk(nbox)), dim3(TPB), 0, st, nbox, T->s_d.p, xb.p, yb.p, nbat.p, npar.p,
                     nchi.p, T->xb.p, T->yb.p, T->nbat.p, T->npar.p, T->nchi.p);
  hipLaunchKernelGGL(k_at_ibox_shift, dim3(nblk(n)), dim3(TPB), 0, st, n, T->s_d.p, T->ibox.p);
  AT_TRY(grow(T->nl, L + 2, 0, st));
  EB_HIP_CHECK(hipMemcpyAsync(T->nl.p, nl.data() + 1, sizeof(int) * (L + 1), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_at_nl_shift, dim3(1), dim3(AT_MAXLEV + 4), 0, st, L + 1, T->s_d.p, T->nl.p);
  EB_HIP_CHECK(hipMemcpyAsync(nl.data() + 1, T->nl.p, sizeof(int) * (L + 1), hipMemcpyDeviceToHost, st));
  EB_HIP_CHECK(hipStreamSynchronize(st));
  nbox = nb;

  // ---- DAINDB
  AT_TRY(grow(T->indb, nbox + 1, 0, st));
  AT_TRY(grow(T->s_e, 2, 0, st));
  EB_HIP_CHECK(hipMemsetAsync(T->s_e.p, 0, sizeof(int) * 2, st));
  hipLaunchKernelGGL(k_at_indb, dim3(nblk(nbox)), dim3(TPB), 0, st, nbox, nlev, T->nl.p, T->nbat.p, T->nchi.p, napb,
                     T->indb.p, T->s_e.p);

  // ---- DAIJAT
  AT_TRY(grow(T->iat, nbox + 1, 0, st));
  AT_TRY(grow(T->s_c, nbox + 1, 0, st));
  AT_TRY(grow(T->s_d, nbox + 2, 0, st));
  hipLaunchKernelGGL(k_at_iat_weight, dim3(nblk(nbox)), dim3(TPB), 0, st, nbox, T->indb.p, T->nbat.p, T->s_c.p);
  AT_TRY(scan(T, T->s_c.p, T->s_d.p, nbox, true));
  int jatlen = 0;
  AT_TRY(read_int(T, T->s_d.p + nbox, &jatlen));
  AT_TRY(grow(T->jat, jatlen + 1, 0, st));
  hipLaunchKernelGGL(k_at_iat, dim3(nblk(nbox)), dim3(TPB), 0, st, nbox, T->indb.p, T->nbat.p, T->s_d.p, T->iat.p,
                     T->jat.p);
  AT_TRY(grow(T->s_k0, n, 0, st));
  AT_TRY(grow(T->s_k1, n, 0, st));
  AT_TRY(grow(T->s_v0, n, 0, st));
  AT_TRY(grow(T->s_v1, n, 0, st));
  hipLaunchKernelGGL(k_at_keys, dim3(nblk(n)), dim3(TPB), 0, st, n, T->ibox.p, T->s_k0.p, T->s_v0.p);
  AT_TRY(grow(T->s_tmp, eb::radix_sort_temp_ints(n) + 1, 0, st));
  AT_TRY(eb::radix_sort_pairs(T->s_k0.p, T->s_v0.p, T->s_k1.p, T->s_v1.p, n, bits_for(nbox + 1), T->s_tmp.p, st));
  AT_TRY(grow(T->s_f, nbox + 1, 0, st));
  hipLaunchKernelGGL(k_at_first, dim3(nblk(n)), dim3(TPB), 0, st, n, T->s_k0.p, T->s_f.p);
  hipLaunchKernelGGL(k_at_jat, dim3(nblk(n)), dim3(TPB), 0, st, n, T->s_k0.p, T->s_v0.p, T->s_f.p, T->iat.p,
                     T->jat.p, T->s_e.p + 1);
  EB_HIP_CHECK(hipGetLastError());

  // ---- DALI12
  AT_TRY(grow(T->l1, nbox + 2, 0, st));
  AT_TRY(grow(T->l2, nbox + 2, 0, st));
  AT_TRY(grow(T->jl, (size_t)(LI_MAX1 + LI_MAX2) * nbox + 16, 0, st));
  hipLaunchKernelGGL(k_at_in12, dim3(1), dim3(64), 0, st, nl[1], nl[2] - 1, T->l1.p, T->l2.p, T->jl.p);
  {
    double d = 0.5 * 1.25 * 128.0;
    for (int lev = 2; lev <= L - 1; ++lev) {
      d = d / 2;
      const int lo = nl[lev], m = nl[lev + 1] - nl[lev];
      if (m <= 0) continue;
      AT_TRY(grow(T->s_a, m + 1, 0, st));
      AT_TRY(grow(T->s_b, m + 1, 0, st));
      hipLaunchKernelGGL(k_at_li12_count, dim3(nblk(m)), dim3(TPB), 0, st, lo, m, d, T->xb.p, T->yb.p, T->npar.p,
                         T->nchi.p, T->l1.p, T->l2.p, T->jl.p, T->s_a.p, T->s_e.p + 1);
      AT_TRY(scan(T, T->s_a.p, T->s_b.p, m, false));
      hipLaunchKernelGGL(k_at_li12_write, dim3(nblk(m)), dim3(TPB), 0, st, lo, m, d, T->xb.p, T->yb.p, T->npar.p,
                         T->nchi.p, T->l1.p, T->l2.p, T->jl.p, T->s_b.p);
      EB_HIP_CHECK(hipGetLastError());
    }
  }
  int flags[2] = {0, 0};
  EB_HIP_CHECK(hipMemcpyAsync(flags, T->s_e.p, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
  int l1end = 0;
  AT_TRY(read_int(T, T->l1.p + nbox, &l1end));
  if (flags[1]) {
    eb_set_error("adaptive tree: inconsistent tables (flags %d: 1 = a list-1/list-2 count above 27/8, "
                 "2 = a point outside every childless box)", flags[1]);
    return EB_ERR_ARG;
  }
  T->maxp = flags[0];
  T->len = l1end - 1;

  // ---- DALI34
  AT_TRY(grow(T->s_d1, nbox + 1, 0, st));
  AT_TRY(grow(T->s_a, nbox + 1, 0, st));
  AT_TRY(grow(T->s_b, nbox + 2, 0, st));
  hipLaunchKernelGGL(k_at_li34_init, dim3(nblk(nbox)), dim3(TPB), 0, st, nbox, T->indb.p, T->l1.p, T->l2.p, T->jl.p,
                     T->xb.p, T->yb.p, T->s_a.p, T->s_d1.p);
  AT_TRY(scan(T, T->s_a.p, T->s_b.p, nbox, true));
  int nf = 0;
  AT_TRY(read_int(T, T->s_b.p + nbox, &nf));
  DBuf<int> fo, fn, nfo, nfn, f3, f4, fc, e3, e4, ec, o3o, o3n, o4o, o4n;
  struct Guard2 {  // frontier and list staging buffers
    hipStream_t st;
    std::vector<DBuf<int>*> v;
    ~Guard2() {
      (void)hipStreamSynchronize(st);
      for (auto* b : v) release(*b);
    }
  } guard2{st, {&fo, &fn, &nfo, &nfn, &f3, &f4, &fc, &e3, &e4, &ec, &o3o, &o3n, &o4o, &o4n}};
  AT_TRY(grow(fo, nf + 1, 0, st));
  AT_TRY(grow(fn, nf + 1, 0, st));
  hipLaunchKernelGGL(k_at_li34_front, dim3(nblk(nbox)), dim3(TPB), 0, st, nbox, T->s_a.p, T->s_b.p, T->l2.p, T->jl.p,
                     fo.p, fn.p);
  int len3 = 0, len4 = 0;
  for (int iloop = 1; nf > 0; ++iloop) {
    // DASCAN's bound D = D1 * (0.5 + 1.25 * 0.5**ILOOP), the factor in REAL*4
    volatile float p = ldexpf(1.0f, -iloop);
    volatile float q = 1.25f * p;
    volatile float fac = 0.5f + q;
    AT_TRY(grow(f3, nf + 1, 0, st));
    AT_TRY(grow(f4, nf + 1, 0, st));
    AT_TRY(grow(fc, nf + 1, 0, st));
    AT_TRY(grow(e3, nf + 1, 0, st));
    AT_TRY(grow(e4, nf + 1, 0, st));
    AT_TRY(grow(ec, nf + 1, 0, st));
    hipLaunchKernelGGL(k_at_li34_classify, dim3(nblk(nf)), dim3(TPB), 0, st, nf, fo.p, fn.p, T->xb.p, T->yb.p,
                       T->indb.p, T->nchi.p, T->npar.p, T->s_d1.p, (double)fac, f3.p, f4.p, fc.p);
    AT_TRY(scan(T, f3.p, e3.p, nf, true));
    AT_TRY(scan(T, f4.p, e4.p, nf, true));
    AT_TRY(scan(T, fc.p, ec.p, nf, true));
    int t3 = 0, t4 = 0, tc = 0;
    EB_HIP_CHECK(hipMemcpyAsync(&t3, e3.p + nf, sizeof(int), hipMemcpyDeviceToHost, st));
    EB_HIP_CHECK(hipMemcpyAsync(&t4, e4.p + nf, sizeof(int), hipMemcpyDeviceToHost, st));
    EB_HIP_CHECK(hipMemcpyAsync(&tc, ec.p + nf, sizeof(int), hipMemcpyDeviceToHost, st));
    EB_HIP_CHECK(hipStreamSynchronize(st));
    AT_TRY(grow(o3o, (size_t)len3 + t3 + 1, len3, st));
    AT_TRY(grow(o3n, (size_t)len3 + t3 + 1, len3, st));
    AT_TRY(grow(o4o, (size_t)len4 + t4 + 1, len4, st));
    AT_TRY(grow(o4n, (size_t)len4 + t4 + 1, len4, st));
    AT_TRY(grow(nfo, tc + 1, 0, st));
    AT_TRY(grow(nfn, tc + 1, 0, st));
    hipLaunchKernelGGL(k_at_li34_emit, dim3(nblk(nf)), dim3(TPB), 0, st, nf, fo.p, fn.p, f3.p, f4.p, fc.p, e3.p,
                       e4.p, ec.p, len3, len4, o3o.p, o3n.p, o4o.p, o4n.p, nfo.p, nfn.p, T->nchi.p);
    EB_HIP_CHECK(hipGetLastError());
    len3 += t3;
    len4 += t4;
    std::swap(fo, nfo);
    std::swap(fn, nfn);
    nf = tc;
  }
  T->nbox = nbox;
  AT_TRY(group_by_owner(T, len3, o3o.p, o3n.p, T->l3, T->jl3));
  AT_TRY(group_by_owner(T, len4, o4o.p, o4n.p, T->l4, T->jl4));
  EB_HIP_CHECK(hipStreamSynchronize(st));
  T->nlev = nlev;
  T->len3 = len3;
  T->len4 = len4;
  T->jatlen = jatlen;
  return EB_OK;
}

}  // namespace

extern "C" int eb_atree_build(eb_atree* t, int32_t nat, const double* x, const double* y, int32_t napb, int32_t lmx,
                              double eps7, int32_t* ierr, void* stream) {
  return eb_atree_build_bounded(t, nat, x, y, napb, lmx, eps7, -1, ierr, stream);
}

extern "C" int eb_atree_build_bounded(eb_atree* t, int32_t nat, const double* x, const double* y, int32_t napb,
                                      int32_t lmx, double eps7, int64_t nboxmx, int32_t* ierr, void* stream) {
  if (nboxmx < 0) nboxmx = INT64_MAX;
  int32_t dummy[3];
  if (!ierr) ierr = dummy;
  ierr[0] = ierr[1] = ierr[2] = 0;
  if (!t) {
    eb_set_error("eb_atree_build: null handle pointer");
    return EB_ERR_ARG;
  }
  *t = nullptr;
  if (nat > 0 && (!x || !y)) {
    eb_set_error("eb_atree_build: null coordinate pointer");
    return EB_ERR_ARG;
  }
  if (napb < 1) {
    eb_set_error("eb_atree_build: NAPB = %d", napb);
    return EB_ERR_ARG;
  }
  if (nat <= napb) {  // dapif1.f:711-717
    ierr[0] = EB_ATREE_FEW_POINTS;
    ierr[1] = nat;
    ierr[2] = napb;
    return EB_ATREE_FEW_POINTS;
  }
  eb_atree_s* T = new eb_atree_s();
  T->nat = nat;
  T->napb = napb;
  T->st = (hipStream_t)stream;
  const int rc = build(T, x, y, std::min<int>(lmx, AT_MAXLEV), eps7, (long long)nboxmx, ierr);
  if (rc != EB_OK) {
    (void)hipStreamSynchronize(T->st);
    delete T;
    return rc;
  }
  T->free_scratch();
  *t = T;
  return EB_OK;
}

extern "C" int eb_atree_counts(eb_atree t, int32_t* counts) {
  if (!t || !counts) {
    eb_set_error("eb_atree_counts: null argument");
    return EB_ERR_ARG;
  }
  const int32_t c[8] = {t->nbox, t->nlev, t->len, t->len3, t->len4, t->maxp, t->jatlen, t->nat};
  for (int i = 0; i < 8; ++i) counts[i] = c[i];
  return EB_OK;
}

extern "C" int eb_atree_tables(eb_atree t, eb_atree_view* v) {
  if (!t || !v) {
    eb_set_error("eb_atree_tables: null argument");
    return EB_ERR_ARG;
  }
  *v = eb_atree_view{t->xs.p, t->ys.p, t->geom.p, t->xb.p, t->yb.p, t->nl.p, t->ibox.p, t->nbat.p,
                     t->npar.p, t->nchi.p, t->indb.p, t->iat.p, t->jat.p, t->l1.p, t->l2.p, t->jl.p,
                     t->l3.p, t->l4.p, t->jl3.p, t->jl4.p};
  return EB_OK;
}

extern "C" int eb_atree_destroy(eb_atree t) {
  if (t) {
    (void)hipStreamSynchronize(t->st);
    delete t;
  }
  return EB_OK;
}
