// eb_adjoint.hip -- the layer potentials of F as O(N^2) pair sums with their vector-Jacobian
// products, and F's Laplace and Stokes operators with their transposes (include/eb_adjoint.h).
//
// Every pair sum runs on a two-dimensional grid: blockIdx.x picks 256 output points (one per
// thread), blockIdx.y a chunk of the input points, staged 256 at a time through LDS.  Each
// block writes its chunk's partial result; a second kernel adds the chunks in chunk order, so
// results do not depend on scheduling (finite differences of the forward need that).  The
// number of chunks is chosen so that a launch has about 2048 blocks at the sizes of F (a few
// thousand points), enough to fill the 256 CUs.
//
// Adjoint convention: for a real loss L and complex v, vbar = dL/dRe v + i dL/dIm v.  A term
// a dv + b dconj(v) of a complex output G with adjoint Gbar contributes Gbar conj(a) + conj(Gbar) b
// to vbar.
#include <math.h>

#include <algorithm>
#include <vector>

#include "../../include/eb_adjoint.h"
#include "../../include/eb_cadap.h"
#include "../../include/eb_gmres.h"
#include "eb_internal.h"

using namespace eb;

namespace {

constexpr double PI = 3.14159265358979323846;
constexpr int PT = 256;          // threads of a pair block, and points per LDS tile
constexpr int EWT = 256;         // threads of the pointwise kernels
constexpr int TARGET_BLOCKS = 2048;

struct Pt {
  double2 z, a, b;
};

__device__ __forceinline__ double2 c2(double x, double y) { return make_double2(x, y); }
__device__ __forceinline__ double2 cadd3(double2 a, double2 b, double2 c) {
  return make_double2(a.x + b.x + c.x, a.y + b.y + c.y);
}

// ---- pair operators ------------------------------------------------------------------------
// load(l, p): the payload of input point l; ostate(o, p): that of output point o;
// pair(o, l, acc): one pair; finish(o, acc, r): the chunk's NOUT outputs (linear in acc).

// g_i = sum_j c_j / (t_i - s_j): outputs t, inputs (s, c)
struct CauchyFwd {
  static constexpr int NACC = 1, NOUT = 1;
  const double2 *t, *s, *c;
  __device__ void load(int l, Pt& p) const { p.z = s[l]; p.a = c[l]; }
  __device__ void ostate(int o, Pt& p) const { p.z = t[o]; }
  __device__ void pair(const Pt& o, const Pt& l, double2* acc) const {
    const double2 w = crcp(csub(o.z, l.z));
    cfma(acc[0], l.a, w);
  }
  __device__ void finish(const Pt&, const double2* acc, double2* r) const { r[0] = acc[0]; }
};

// tbar_i = -gbar_i conj(sum_j c_j / (t_i - s_j)^2)
struct CauchyVjpT {
  static constexpr int NACC = 1, NOUT = 1;
  const double2 *t, *s, *c, *gbar;
  __device__ void load(int l, Pt& p) const { p.z = s[l]; p.a = c[l]; }
  __device__ void ostate(int o, Pt& p) const { p.z = t[o]; p.a = gbar[o]; }
  __device__ void pair(const Pt& o, const Pt& l, double2* acc) const {
    const double2 w = crcp(csub(o.z, l.z));
    cfma(acc[0], l.a, cmul(w, w));
  }
  __device__ void finish(const Pt& o, const double2* acc, double2* r) const {
    const double2 v = cmul(o.a, cconj(acc[0]));
    r[0] = c2(-v.x, -v.y);
  }
};

// outputs s_j: sbar_j = conj(c_j sum_i conj(gbar_i) / (t_i - s_j)^2),
//              cbar_j = conj(sum_i conj(gbar_i) / (t_i - s_j))
struct CauchyVjpS {
  static constexpr int NACC = 2, NOUT = 2;
  const double2 *t, *s, *c, *gbar;
  __device__ void load(int l, Pt& p) const { p.z = t[l]; p.a = cconj(gbar[l]); }
  __device__ void ostate(int o, Pt& p) const { p.z = s[o]; p.a = c[o]; }
  __device__ void pair(const Pt& o, const Pt& l, double2* acc) const {
    const double2 w = crcp(csub(l.z, o.z));
    const double2 gw = cmul(l.a, w);
    acc[0] = cadd(acc[0], gw);
    cfma(acc[1], gw, w);
  }
  __device__ void finish(const Pt& o, const double2* acc, double2* r) const {
    r[0] = cconj(cmul(o.a, acc[1]));
    r[1] = cconj(acc[0]);
  }
};

// gw_i = sum_j qa_j / r + s2 (conj(qb_j) / conj(r) - conj(qa_j) r / conj(r)^2), r = z_j - z_i
struct StokesFwd {
  static constexpr int NACC = 2, NOUT = 1;
  const double2 *z, *qa, *qb;
  double s2;
  __device__ void load(int l, Pt& p) const { p.z = z[l]; p.a = qa[l]; p.b = qb[l]; }
  __device__ void ostate(int o, Pt& p) const { p.z = z[o]; }
  __device__ void pair(const Pt& o, const Pt& l, double2* acc) const {
    p2p_pair<false, false>(o.z, l.z, l.a, l.b, false, acc[0], acc[1]);
  }
  __device__ void finish(const Pt&, const double2* acc, double2* r) const {
    r[0] = c2(acc[0].x + s2 * acc[1].x, acc[0].y + s2 * acc[1].y);
  }
};

// The pair term f(r) above has df = a dr + b dconj(r) with
//   a = -qa / r^2 - s2 conj(qa) / conj(r)^2,   b = s2 (-conj(qb) / conj(r)^2 + 2 conj(qa) r / conj(r)^3).
// Target side: zbar_i = -(gbar_i conj(sum_j a) + conj(gbar_i) sum_j b).
struct StokesVjpT {
  static constexpr int NACC = 2, NOUT = 1;
  const double2 *z, *qa, *qb, *gbar;
  double s2;
  __device__ void load(int l, Pt& p) const { p.z = z[l]; p.a = qa[l]; p.b = qb[l]; }
  __device__ void ostate(int o, Pt& p) const { p.z = z[o]; p.a = gbar[o]; }
  __device__ void pair(const Pt& o, const Pt& l, double2* acc) const {
    const double2 r = csub(l.z, o.z);
    const double2 w = crcp(r);
    const double2 cw = cconj(w);
    const double2 w2 = cmul(w, w), cw2 = cmul(cw, cw);
    const double2 cqa = cconj(l.a), cqb = cconj(l.b);
    // a = -qa w^2 - s2 conj(qa) cw^2
    const double2 t1 = cmul(l.a, w2), t2 = cmul(cqa, cw2);
    acc[0] = c2(acc[0].x - t1.x - s2 * t2.x, acc[0].y - t1.y - s2 * t2.y);
    // b = s2 (-conj(qb) cw^2 + 2 conj(qa) r cw^3)
    const double2 t3 = cmul(cqb, cw2);
    const double2 t4 = cmul(cmul(t2, cw), r);
    acc[1] = c2(acc[1].x + s2 * (2.0 * t4.x - t3.x), acc[1].y + s2 * (2.0 * t4.y - t3.y));
  }
  __device__ void finish(const Pt& o, const double2* acc, double2* r) const {
    const double2 u = cadd(cmul(o.a, cconj(acc[0])), cmul(cconj(o.a), acc[1]));
    r[0] = c2(-u.x, -u.y);
  }
};

// Source side, outputs j (r = z_j - z_i, w = 1/r, G = gbar_i):
//   Q1 = sum G w, Q2 = sum G w^2, Q3 = sum G conj(w), Q4 = sum G conj(w)^2,
//   Q5 = sum conj(G) r conj(w)^2, Q6 = sum conj(G) r conj(w)^3;
//   zbar_j = -conj(qa) Q4 - s2 qa Q2 + s2 (-conj(qb) conj(Q2) + 2 conj(qa) Q6)
//   qabar_j = Q3 - s2 Q5,  qbbar_j = s2 conj(Q1)
struct StokesVjpS {
  static constexpr int NACC = 6, NOUT = 3;
  const double2 *z, *qa, *qb, *gbar;
  double s2;
  __device__ void load(int l, Pt& p) const { p.z = z[l]; p.a = gbar[l]; }
  __device__ void ostate(int o, Pt& p) const { p.z = z[o]; p.a = qa[o]; p.b = qb[o]; }
  __device__ void pair(const Pt& o, const Pt& l, double2* acc) const {
    const double2 r = csub(o.z, l.z);
    const double2 w = crcp(r);
    const double2 cw = cconj(w);
    const double2 G = l.a, cG = cconj(l.a);
    const double2 Gw = cmul(G, w), Gcw = cmul(G, cw);
    acc[0] = cadd(acc[0], Gw);
    cfma(acc[1], Gw, w);
    acc[2] = cadd(acc[2], Gcw);
    cfma(acc[3], Gcw, cw);
    const double2 t = cmul(cmul(cG, r), cmul(cw, cw));
    acc[4] = cadd(acc[4], t);
    cfma(acc[5], t, cw);
  }
  __device__ void finish(const Pt& o, const double2* acc, double2* r) const {
    const double2 qa_ = o.a, cqa = cconj(o.a), cqb = cconj(o.b);
    const double2 a1 = cmul(cqa, acc[3]);
    const double2 a2 = cmul(qa_, acc[1]);
    const double2 a3 = cmul(cqb, cconj(acc[1]));
    const double2 a4 = cmul(cqa, acc[5]);
    r[0] = c2(-a1.x + s2 * (-a2.x - a3.x + 2.0 * a4.x), -a1.y + s2 * (-a2.y - a3.y + 2.0 * a4.y));
    r[1] = c2(acc[2].x - s2 * acc[4].x, acc[2].y - s2 * acc[4].y);
    r[2] = c2(s2 * acc[0].x, -s2 * acc[0].y);
  }
};

// The transpose of sigma -> GW(qa(sigma), qb(sigma)) with qa = alpha sigma,
// qb = -alpha conj(sigma) + gamma sigma (alpha = dz h / (2 pi i), gamma = conj(dz) h / (2 pi i)):
//   mu_j = conj(alpha_j) (S1 - s2 S4) - s2 alpha_j S2 + s2 conj(gamma_j) conj(S2),
//   S1 = sum_i lam_i conj(w), S2 = sum_i lam_i w, S4 = sum_i conj(lam_i) r conj(w)^2,  r = z_j - z_i
struct StokesT {
  static constexpr int NACC = 3, NOUT = 1;
  const double2 *z, *lam, *alpha, *gamma;
  double s2;
  __device__ void load(int l, Pt& p) const { p.z = z[l]; p.a = lam[l]; }
  __device__ void ostate(int o, Pt& p) const { p.z = z[o]; p.a = alpha[o]; p.b = gamma[o]; }
  __device__ void pair(const Pt& o, const Pt& l, double2* acc) const {
    const double2 r = csub(o.z, l.z);
    const double2 w = crcp(r);
    const double2 cw = cconj(w);
    cfma(acc[0], l.a, cw);
    cfma(acc[1], l.a, w);
    cfma(acc[2], cmul(cconj(l.a), r), cmul(cw, cw));
  }
  __device__ void finish(const Pt& o, const double2* acc, double2* r) const {
    const double2 s14 = c2(acc[0].x - s2 * acc[2].x, acc[0].y - s2 * acc[2].y);
    const double2 u1 = cmul(cconj(o.a), s14);
    const double2 u2 = cmul(o.a, acc[1]);
    const double2 u3 = cmul(cconj(o.b), cconj(acc[1]));
    r[0] = c2(u1.x + s2 * (u3.x - u2.x), u1.y + s2 * (u3.y - u2.y));
  }
};

template <class Op, bool SELF>
__global__ __launch_bounds__(PT) void k_pairs(Op op, int no, int nl, int chunk, double2* __restrict__ part) {
  __shared__ Pt sl[PT];
  const int o = blockIdx.x * PT + threadIdx.x;
  const int l0 = blockIdx.y * chunk;
  const int l1 = min(nl, l0 + chunk);
  Pt po;
  op.ostate(o < no ? o : no - 1, po);
  double2 acc[Op::NACC];
#pragma unroll
  for (int k = 0; k < Op::NACC; ++k) acc[k] = make_double2(0.0, 0.0);
  for (int base = l0; base < l1; base += PT) {
    const int l = base + (int)threadIdx.x;
    if (l < l1) op.load(l, sl[threadIdx.x]);
    __syncthreads();
    const int cnt = min(PT, l1 - base);
    for (int q = 0; q < cnt; ++q) {
      if (SELF && base + q == o) continue;
      op.pair(po, sl[q], acc);
    }
    __syncthreads();
  }
  if (o < no) {
    double2 r[Op::NOUT];
    op.finish(po, acc, r);
#pragma unroll
    for (int k = 0; k < Op::NOUT; ++k) part[((size_t)k * gridDim.y + blockIdx.y) * no + o] = r[k];
  }
}

// out_k[o] = sum over chunks c (in order) of part[k][c][o]
__global__ void k_chunk_sum(int no, int nch, int nout, const double2* __restrict__ part, double2* __restrict__ o0,
                            double2* __restrict__ o1, double2* __restrict__ o2) {
  const int o = blockIdx.x * EWT + threadIdx.x;
  if (o >= no) return;
  double2* outs[3] = {o0, o1, o2};
  for (int k = 0; k < nout; ++k) {
    if (!outs[k]) continue;
    double2 s = make_double2(0.0, 0.0);
    for (int c = 0; c < nch; ++c) s = cadd(s, part[((size_t)k * nch + c) * no + o]);
    outs[k][o] = s;
  }
}

__global__ void k_add_to(int n, const double2* __restrict__ a, double2* __restrict__ y) {
  const int i = blockIdx.x * EWT + threadIdx.x;
  if (i < n) y[i] = cadd(y[i], a[i]);
}

// ---- pointwise parts of the operators -------------------------------------------------------

__global__ void k_body_of_a(const int* __restrict__ off, int nbody, int* __restrict__ body_of) {
  const int k = blockIdx.x;
  for (int i = off[k] + threadIdx.x; i < off[k + 1]; i += blockDim.x) body_of[i] = k;
}

// per-point coefficients of the operators from the geometry
__global__ void k_geom(int n, const int* __restrict__ body_of, const double* __restrict__ hb,
                       const double2* __restrict__ z, const double2* __restrict__ dz, const double* __restrict__ s,
                       const double* __restrict__ kap, const double2* __restrict__ zk, double2* __restrict__ wdz,
                       double* __restrict__ lapd, double* __restrict__ c0, double2* __restrict__ zn2,
                       double2* __restrict__ e, double2* __restrict__ alpha, double2* __restrict__ gamma) {
  const int i = blockIdx.x * EWT + threadIdx.x;
  if (i >= n) return;
  const int b = body_of[i];
  const double h = hb[b];
  const double2 d = dz[i];
  wdz[i] = c2(d.x * h, d.y * h);
  lapd[i] = 1.0 + h * kap[i] * s[i] / (2.0 * PI) - 2.0 * h * s[i];
  c0[i] = h * s[i] * kap[i] / (2.0 * PI);
  const double2 zn = c2(d.y / s[i], -d.x / s[i]);  // -i dz / s
  zn2[i] = cmul(zn, zn);
  const double2 q = csub(z[i], zk[b]);
  e[i] = cmul(d, crcp(cmul(q, q)));
  // dz h / (2 pi i) = -i dz h / (2 pi)
  const double f = h / (2.0 * PI);
  alpha[i] = c2(d.y * f, -d.x * f);
  gamma[i] = c2(-d.y * f, -d.x * f);  // -i conj(dz) h / (2 pi)
}

__global__ void k_lap_charges(int n, const double* __restrict__ u, const double2* __restrict__ wdz,
                              double2* __restrict__ c) {
  const int i = blockIdx.x * EWT + threadIdx.x;
  if (i < n) c[i] = c2(u[i] * wdz[i].x, u[i] * wdz[i].y);
}

__global__ void k_real_c(int n, const double* __restrict__ u, double2* __restrict__ c) {
  const int i = blockIdx.x * EWT + threadIdx.x;
  if (i < n) c[i] = c2(u[i], 0.0);
}

// y = d u - Im(cf) / pi  (transpose: y = d u + Im(wdz cf) / pi)
__global__ void k_lap_out(int n, int transpose, const double* __restrict__ u, const double* __restrict__ lapd,
                          const double2* __restrict__ wdz, const double2* __restrict__ cf, double* __restrict__ y) {
  const int i = blockIdx.x * EWT + threadIdx.x;
  if (i >= n) return;
  if (transpose)
    y[i] = lapd[i] * u[i] + cmul(wdz[i], cf[i]).y / PI;
  else
    y[i] = lapd[i] * u[i] - cf[i].y / PI;
}

__global__ void k_st_charges(int n, const double2* __restrict__ sg, const double2* __restrict__ alpha,
                             const double2* __restrict__ gamma, double2* __restrict__ qa, double2* __restrict__ qb) {
  const int i = blockIdx.x * EWT + threadIdx.x;
  if (i >= n) return;
  const double2 a = alpha[i], v = sg[i];
  qa[i] = cmul(a, v);
  const double2 t = cmul(a, cconj(v));
  const double2 g = cmul(gamma[i], v);
  qb[i] = c2(g.x - t.x, g.y - t.y);
}

// per body: forward  s1 = h sum sigma, s2 = (2h/pi) sum Re(sigma e);
//           transpose s1 = h sum lam,  s2 = sum Re(conj(lam) (i/2) z)
__global__ __launch_bounds__(256) void k_body_sums(const int* __restrict__ off, const double* __restrict__ hb,
                                                   int transpose, const double2* __restrict__ v,
                                                   const double2* __restrict__ z, const double2* __restrict__ e,
                                                   double2* __restrict__ s1, double* __restrict__ s2) {
  __shared__ double red[3][256];
  const int b = blockIdx.x;
  double a = 0, c = 0, p = 0;
  for (int i = off[b] + threadIdx.x; i < off[b + 1]; i += blockDim.x) {
    const double2 x = v[i];
    a += x.x;
    c += x.y;
    if (transpose)
      p += -0.5 * (x.x * z[i].y - x.y * z[i].x);  // Re(conj(x) i z / 2) = -Im(conj(x) z) / 2
    else
      p += x.x * e[i].x - x.y * e[i].y;  // Re(x e)
  }
  red[0][threadIdx.x] = a;
  red[1][threadIdx.x] = c;
  red[2][threadIdx.x] = p;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double h = hb[b];
    s1[b] = c2(h * red[0][0], h * red[1][0]);
    s2[b] = transpose ? red[2][0] : 2.0 * h * red[2][0] / PI;
  }
}

// forward:   y = sg + za_b + gw + c0 (sg - conj(sg) zn2) + (i/2) pk_b z
// transpose: y = sg + za_b + gw + c0 (sg - conj(sg) zn2) + (2 h / pi) rho_b conj(e)
__global__ void k_st_out(int n, int transpose, const int* __restrict__ body_of, const double* __restrict__ hb,
                         const double2* __restrict__ sg, const double2* __restrict__ za, const double* __restrict__ pk,
                         const double2* __restrict__ gw, const double* __restrict__ c0, const double2* __restrict__ zn2,
                         const double2* __restrict__ z, const double2* __restrict__ e, double2* __restrict__ y) {
  const int i = blockIdx.x * EWT + threadIdx.x;
  if (i >= n) return;
  const int b = body_of[i];
  const double2 v = sg[i];
  const double2 t = cmul(cconj(v), zn2[i]);
  const double2 self = c2(c0[i] * (v.x - t.x), c0[i] * (v.y - t.y));
  double2 last;
  if (transpose) {
    const double f = 2.0 * hb[b] * pk[b] / PI;
    last = c2(f * e[i].x, -f * e[i].y);
  } else {
    const double a = 0.5 * pk[b];
    last = c2(-a * z[i].y, a * z[i].x);
  }
  y[i] = cadd(cadd3(v, za[b], gw[i]), cadd(self, last));
}

// coordinates for the FMM plans, and their unit weights
__global__ void k_split_xy(int n, const double2* __restrict__ z, double* __restrict__ x, double* __restrict__ y,
                           double2* __restrict__ ones) {
  const int i = blockIdx.x * EWT + threadIdx.x;
  if (i >= n) return;
  x[i] = z[i].x;
  y[i] = z[i].y;
  ones[i] = c2(1.0, 0.0);
}

}  // namespace

struct eb_adj_s {
  int nbody = 0, n = 0, maxl = 100;
  hipStream_t st = nullptr;
  std::vector<int> off;
  int *d_off = nullptr, *body_of = nullptr;
  double* hb = nullptr;
  bool have_geom = false;
  double2 *z = nullptr, *dz = nullptr, *zk = nullptr, *wdz = nullptr, *zn2 = nullptr, *e = nullptr, *alpha = nullptr,
          *gamma = nullptr;
  double *s = nullptr, *kap = nullptr, *lapd = nullptr, *c0 = nullptr;
  double2 *c = nullptr, *cf = nullptr, *qa = nullptr, *qb = nullptr, *gw = nullptr, *bs1 = nullptr;
  double* bs2 = nullptr;
  double2* part = nullptr;
  size_t part_cap = 0;
  eb_gmres glap = nullptr, gst = nullptr;
  // EB_MODE_FMM: the operators' layer potentials on FMM plans over the same points (the
  // Stokes kernel on CADAPDR's, the Laplace Cauchy sum on the Hilbert product with unit
  // weights), their transposes by eb_fmm_apply_transpose
  int mode = EB_MODE_AUTO;
  eb_fmm_plan pst = nullptr, plap = nullptr;
  double *px = nullptr, *py = nullptr;
  double2* ones = nullptr;
};

namespace {

template <typename T>
int dalloc_a(T** p, size_t count) {
  *p = nullptr;
  if (hipMalloc((void**)p, sizeof(T) * (count ? count : 1)) != hipSuccess) {
    eb_set_error("hipMalloc(%zu bytes) failed", sizeof(T) * count);
    return EB_ERR_NOMEM;
  }
  return EB_OK;
}

void free_adj(eb_adj A) {
  if (!A) return;
  if (A->st) (void)hipStreamSynchronize(A->st);
  if (A->glap) eb_gmres_destroy(A->glap);
  if (A->gst) eb_gmres_destroy(A->gst);
  if (A->pst) eb_fmm_plan_destroy(A->pst);
  if (A->plap) eb_fmm_plan_destroy(A->plap);
  void* p[] = {A->d_off, A->body_of, A->hb, A->z, A->dz, A->zk, A->wdz, A->zn2, A->e, A->alpha, A->gamma, A->s, A->kap,
               A->lapd, A->c0, A->c, A->cf, A->qa, A->qb, A->gw, A->bs1, A->bs2, A->part, A->px, A->py, A->ones};
  for (void* q : p)
    if (q) (void)hipFree(q);
  delete A;
}

// chunking of nl input points for no outputs: about TARGET_BLOCKS blocks, chunks of whole tiles
void chunking(int no, int nl, int* nch, int* chunk) {
  const int tiles = (no + PT - 1) / PT;
  const int ltiles = (nl + PT - 1) / PT;
  int c = std::max(1, std::min(ltiles, (TARGET_BLOCKS + tiles - 1) / tiles));
  const int per = (ltiles + c - 1) / c;  // tiles per chunk
  *chunk = per * PT;
  *nch = (nl + *chunk - 1) / *chunk;
}

int ensure_part(eb_adj A, size_t count) {
  if (count <= A->part_cap) return EB_OK;
  if (A->part) {
    EB_HIP_CHECK(hipStreamSynchronize(A->st));
    EB_HIP_CHECK(hipFree(A->part));
    A->part = nullptr;
    A->part_cap = 0;
  }
  int rc = dalloc_a(&A->part, count);
  if (rc == EB_OK) A->part_cap = count;
  return rc;
}

template <class Op>
int run_pairs(eb_adj A, const Op& op, bool self, int no, int nl, double2* o0, double2* o1, double2* o2) {
  if (no < 1 || nl < 1) return (eb_set_error("pair sum: empty point set"), EB_ERR_ARG);
  int nch, chunk;
  chunking(no, nl, &nch, &chunk);
  int rc = ensure_part(A, (size_t)Op::NOUT * nch * no);
  if (rc != EB_OK) return rc;
  const dim3 grid((no + PT - 1) / PT, nch);
  if (self)
    hipLaunchKernelGGL((k_pairs<Op, true>), grid, dim3(PT), 0, A->st, op, no, nl, chunk, A->part);
  else
    hipLaunchKernelGGL((k_pairs<Op, false>), grid, dim3(PT), 0, A->st, op, no, nl, chunk, A->part);
  hipLaunchKernelGGL(k_chunk_sum, dim3((no + EWT - 1) / EWT), dim3(EWT), 0, A->st, no, nch, Op::NOUT, A->part, o0, o1,
                     o2);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

inline int ewb(int n) { return (n + EWT - 1) / EWT; }

// the operators run on FMM plans: EB_MODE_FMM, or EB_MODE_AUTO above EB_DIRECT_MAX points (the
// drop-in layer's rule)
bool use_fmm(eb_adj A) { return A->mode == EB_MODE_FMM || (A->mode == EB_MODE_AUTO && A->n > EB_DIRECT_MAX); }

// LEVMAX of the plans: about 16..64 points per box of the finest uniform grid were the points
// spread evenly (4^L >= n / 48), between 3 and 10
int adj_levmax(int n) {
  int L = 3;
  while (L < 10 && ((int64_t)1 << (2 * L)) * 48 < (int64_t)n) ++L;
  return L;
}

// the expansion length of the plans: CADAPMOD's TOL, 1e-13 (electric.f:803), -NINT(log2 TOL)
constexpr int ADJ_NTERMS = 43;

// builds or re-bins the plans for the current geometry (FMM mode only)
int sync_plans(eb_adj A) {
  if (!use_fmm(A) || !A->have_geom) return EB_OK;
  const int n = A->n;
  int rc;
  if (!A->px) {
    if ((rc = dalloc_a(&A->px, n)) || (rc = dalloc_a(&A->py, n)) || (rc = dalloc_a(&A->ones, n))) return rc;
  }
  hipLaunchKernelGGL(k_split_xy, dim3(ewb(n)), dim3(EWT), 0, A->st, n, A->z, A->px, A->py, A->ones);
  EB_HIP_CHECK(hipGetLastError());
  if (!A->pst) {
    const int L = adj_levmax(n);
    if ((rc = eb_fmm_plan_create(&A->pst, n, A->px, A->py, (const double*)A->wdz, L, ADJ_NTERMS, EB_VARIANT_DR,
                                 EB_MODE_FMM, A->st)) != EB_OK)
      return rc;
    return eb_fmm_plan_create(&A->plap, n, A->px, A->py, (const double*)A->ones, L, ADJ_NTERMS, EB_VARIANT_HILBERT,
                              EB_MODE_FMM, A->st);
  }
  if ((rc = eb_fmm_plan_rebuild(A->pst, A->px, A->py, (const double*)A->wdz, A->st)) != EB_OK) return rc;
  return eb_fmm_plan_rebuild(A->plap, A->px, A->py, (const double*)A->ones, A->st);
}

int need_geom(eb_adj A, const char* what) {
  if (!A || !A->have_geom) {
    eb_set_error("%s: no geometry (call eb_adj_set_geometry first)", what);
    return EB_ERR_ARG;
  }
  return EB_OK;
}

int lap_apply_cb(void* ctx, const double* x, double* y, void*) { return eb_adj_laplace_apply((eb_adj)ctx, 0, x, y); }
int lap_apply_t_cb(void* ctx, const double* x, double* y, void*) { return eb_adj_laplace_apply((eb_adj)ctx, 1, x, y); }
int st_apply_cb(void* ctx, const double* x, double* y, void*) { return eb_adj_stokes_apply((eb_adj)ctx, 0, x, y); }
int st_apply_t_cb(void* ctx, const double* x, double* y, void*) { return eb_adj_stokes_apply((eb_adj)ctx, 1, x, y); }

}  // namespace

extern "C" int eb_adj_create(eb_adj* out, int32_t nbody, const int32_t* nd, int32_t maxl, void* stream) {
  if (!out || nbody < 1 || !nd || maxl < 1) return (eb_set_error("eb_adj_create: bad arguments"), EB_ERR_ARG);
  *out = nullptr;
  eb_adj A = new eb_adj_s();
  A->nbody = nbody;
  A->maxl = maxl;
  A->st = (hipStream_t)stream;
  A->off.assign(nbody + 1, 0);
  std::vector<double> h(nbody);
  for (int b = 0; b < nbody; ++b) {
    if (nd[b] < 4) {
      free_adj(A);
      return (eb_set_error("eb_adj_create: body %d has %d points", b, nd[b]), EB_ERR_ARG);
    }
    A->off[b + 1] = A->off[b] + nd[b];
    h[b] = 2.0 * PI / nd[b];
  }
  const int n = A->n = A->off[nbody];
  int rc;
  if ((rc = dalloc_a(&A->d_off, nbody + 1)) || (rc = dalloc_a(&A->body_of, n)) || (rc = dalloc_a(&A->hb, nbody)) ||
      (rc = dalloc_a(&A->z, n)) || (rc = dalloc_a(&A->dz, n)) || (rc = dalloc_a(&A->zk, nbody)) ||
      (rc = dalloc_a(&A->wdz, n)) || (rc = dalloc_a(&A->zn2, n)) || (rc = dalloc_a(&A->e, n)) ||
      (rc = dalloc_a(&A->alpha, n)) || (rc = dalloc_a(&A->gamma, n)) || (rc = dalloc_a(&A->s, n)) ||
      (rc = dalloc_a(&A->kap, n)) || (rc = dalloc_a(&A->lapd, n)) || (rc = dalloc_a(&A->c0, n)) ||
      (rc = dalloc_a(&A->c, n)) || (rc = dalloc_a(&A->cf, n)) || (rc = dalloc_a(&A->qa, n)) ||
      (rc = dalloc_a(&A->qb, n)) || (rc = dalloc_a(&A->gw, n)) || (rc = dalloc_a(&A->bs1, nbody)) ||
      (rc = dalloc_a(&A->bs2, nbody))) {
    free_adj(A);
    return rc;
  }
  if (hipMemcpyAsync(A->d_off, A->off.data(), sizeof(int) * (nbody + 1), hipMemcpyHostToDevice, A->st) != hipSuccess ||
      hipMemcpyAsync(A->hb, h.data(), sizeof(double) * nbody, hipMemcpyHostToDevice, A->st) != hipSuccess) {
    free_adj(A);
    return (eb_set_error("eb_adj_create: upload failed"), EB_ERR_HIP);
  }
  hipLaunchKernelGGL(k_body_of_a, dim3(nbody), dim3(256), 0, A->st, A->d_off, nbody, A->body_of);
  if ((rc = eb_gmres_create(&A->glap, n, maxl, stream)) != EB_OK ||
      (rc = eb_gmres_create(&A->gst, 2 * n, maxl, stream)) != EB_OK) {
    free_adj(A);
    return rc;
  }
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(A->st) != hipSuccess) {
    free_adj(A);
    return (eb_set_error("eb_adj_create: kernel failed"), EB_ERR_HIP);
  }
  *out = A;
  return EB_OK;
}

extern "C" int eb_adj_destroy(eb_adj A) {
  free_adj(A);
  return EB_OK;
}

extern "C" int eb_adj_set_geometry(eb_adj A, const double* z, const double* dz, const double* dsda,
                                   const double* kappa, const double* zk) {
  if (!A || !z || !dz || !dsda || !kappa || !zk) return (eb_set_error("eb_adj_set_geometry: bad arguments"), EB_ERR_ARG);
  const int n = A->n;
  EB_HIP_CHECK(hipMemcpyAsync(A->z, z, sizeof(double2) * n, hipMemcpyDeviceToDevice, A->st));
  EB_HIP_CHECK(hipMemcpyAsync(A->dz, dz, sizeof(double2) * n, hipMemcpyDeviceToDevice, A->st));
  EB_HIP_CHECK(hipMemcpyAsync(A->s, dsda, sizeof(double) * n, hipMemcpyDeviceToDevice, A->st));
  EB_HIP_CHECK(hipMemcpyAsync(A->kap, kappa, sizeof(double) * n, hipMemcpyDeviceToDevice, A->st));
  EB_HIP_CHECK(hipMemcpyAsync(A->zk, zk, sizeof(double2) * A->nbody, hipMemcpyDeviceToDevice, A->st));
  hipLaunchKernelGGL(k_geom, dim3(ewb(n)), dim3(EWT), 0, A->st, n, A->body_of, A->hb, A->z, A->dz, A->s, A->kap, A->zk,
                     A->wdz, A->lapd, A->c0, A->zn2, A->e, A->alpha, A->gamma);
  EB_HIP_CHECK(hipGetLastError());
  A->have_geom = true;
  return sync_plans(A);
}

extern "C" int eb_adj_set_mode(eb_adj A, int32_t mode) {
  if (!A || (mode != EB_MODE_AUTO && mode != EB_MODE_FMM && mode != EB_MODE_DIRECT))
    return (eb_set_error("eb_adj_set_mode: bad arguments"), EB_ERR_ARG);
  A->mode = mode;
  return sync_plans(A);
}

extern "C" int eb_adj_laplace_apply(eb_adj A, int32_t transpose, const double* x, double* y) {
  int rc = need_geom(A, "eb_adj_laplace_apply");
  if (rc != EB_OK) return rc;
  if (!x || !y) return (eb_set_error("eb_adj_laplace_apply: bad arguments"), EB_ERR_ARG);
  const int n = A->n;
  if (transpose)
    hipLaunchKernelGGL(k_real_c, dim3(ewb(n)), dim3(EWT), 0, A->st, n, x, A->c);
  else
    hipLaunchKernelGGL(k_lap_charges, dim3(ewb(n)), dim3(EWT), 0, A->st, n, x, A->wdz, A->c);
  if (use_fmm(A)) {
    // CF_i = sum_j c_j / (z_i - z_j): the Hilbert product with unit weights
    if ((rc = eb_fmm_apply(A->plap, (const double*)A->c, (double*)A->cf, A->st)) != EB_OK) return rc;
  } else {
    CauchyFwd op{A->z, A->z, A->c};
    if ((rc = run_pairs(A, op, true, n, n, A->cf, nullptr, nullptr)) != EB_OK) return rc;
  }
  hipLaunchKernelGGL(k_lap_out, dim3(ewb(n)), dim3(EWT), 0, A->st, n, (int)transpose, x, A->lapd, A->wdz, A->cf, y);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" int eb_adj_stokes_apply(eb_adj A, int32_t transpose, const double* x, double* y) {
  int rc = need_geom(A, "eb_adj_stokes_apply");
  if (rc != EB_OK) return rc;
  if (!x || !y) return (eb_set_error("eb_adj_stokes_apply: bad arguments"), EB_ERR_ARG);
  const int n = A->n;
  const double2* sg = (const double2*)x;
  if (use_fmm(A)) {
    // GW and its transpose on the CADAPDR plan (the plan forms QA, QB from the density itself)
    rc = transpose ? eb_fmm_apply_transpose(A->pst, x, (double*)A->gw, A->st)
                   : eb_fmm_apply(A->pst, x, (double*)A->gw, A->st);
    if (rc != EB_OK) return rc;
  } else if (transpose) {
    StokesT op{A->z, sg, A->alpha, A->gamma, -1.0};
    if ((rc = run_pairs(A, op, true, n, n, A->gw, nullptr, nullptr)) != EB_OK) return rc;
  } else {
    hipLaunchKernelGGL(k_st_charges, dim3(ewb(n)), dim3(EWT), 0, A->st, n, sg, A->alpha, A->gamma, A->qa, A->qb);
    StokesFwd op{A->z, A->qa, A->qb, -1.0};
    if ((rc = run_pairs(A, op, true, n, n, A->gw, nullptr, nullptr)) != EB_OK) return rc;
  }
  hipLaunchKernelGGL(k_body_sums, dim3(A->nbody), dim3(256), 0, A->st, A->d_off, A->hb, (int)transpose, sg, A->z, A->e,
                     A->bs1, A->bs2);
  hipLaunchKernelGGL(k_st_out, dim3(ewb(n)), dim3(EWT), 0, A->st, n, (int)transpose, A->body_of, A->hb, sg, A->bs1,
                     A->bs2, A->gw, A->c0, A->zn2, A->z, A->e, (double2*)y);
  EB_HIP_CHECK(hipGetLastError());
  return EB_OK;
}

extern "C" int eb_adj_laplace_solve(eb_adj A, int32_t transpose, const double* b, double* x, double tol,
                                    int32_t* iter, double* err, int32_t* ierr) {
  int rc = need_geom(A, "eb_adj_laplace_solve");
  if (rc != EB_OK) return rc;
  int32_t it = 0, ie = 0;
  double er = 0.0;
  rc = eb_gmres_solve(A->glap, transpose ? lap_apply_t_cb : lap_apply_cb, A, b, x, tol, -1, &it, &er, &ie);
  if (iter) *iter = it;
  if (err) *err = er;
  if (ierr) *ierr = ie;
  return rc;
}

extern "C" int eb_adj_stokes_solve(eb_adj A, int32_t transpose, const double* b, double* x, double tol,
                                   int32_t* iter, double* err, int32_t* ierr) {
  int rc = need_geom(A, "eb_adj_stokes_solve");
  if (rc != EB_OK) return rc;
  int32_t it = 0, ie = 0;
  double er = 0.0;
  rc = eb_gmres_solve(A->gst, transpose ? st_apply_t_cb : st_apply_cb, A, b, x, tol, -1, &it, &er, &ie);
  if (iter) *iter = it;
  if (err) *err = er;
  if (ierr) *ierr = ie;
  return rc;
}

extern "C" int eb_adj_cauchy(eb_adj A, int32_t nt, const double* t, int32_t ns, const double* s, const double* c,
                             int32_t self, double* g) {
  if (!A || !t || !s || !c || !g || (self && nt != ns)) return (eb_set_error("eb_adj_cauchy: bad arguments"), EB_ERR_ARG);
  CauchyFwd op{(const double2*)t, (const double2*)s, (const double2*)c};
  return run_pairs(A, op, self != 0, nt, ns, (double2*)g, nullptr, nullptr);
}

extern "C" int eb_adj_cauchy_vjp(eb_adj A, int32_t nt, const double* t, int32_t ns, const double* s, const double* c,
                                 int32_t self, const double* gbar, double* tbar, double* sbar, double* cbar) {
  if (!A || !t || !s || !c || !gbar || (self && nt != ns))
    return (eb_set_error("eb_adj_cauchy_vjp: bad arguments"), EB_ERR_ARG);
  int rc;
  if (tbar) {
    CauchyVjpT op{(const double2*)t, (const double2*)s, (const double2*)c, (const double2*)gbar};
    if ((rc = run_pairs(A, op, self != 0, nt, ns, (double2*)tbar, nullptr, nullptr)) != EB_OK) return rc;
  }
  if (sbar || cbar) {
    CauchyVjpS op{(const double2*)t, (const double2*)s, (const double2*)c, (const double2*)gbar};
    if ((rc = run_pairs(A, op, self != 0, ns, nt, (double2*)sbar, (double2*)cbar, nullptr)) != EB_OK) return rc;
  }
  return EB_OK;
}

extern "C" int eb_adj_stokes_pairs(eb_adj A, int32_t n, const double* z, const double* qa, const double* qb, double s2,
                                   double* gw) {
  if (!A || !z || !qa || !qb || !gw) return (eb_set_error("eb_adj_stokes_pairs: bad arguments"), EB_ERR_ARG);
  StokesFwd op{(const double2*)z, (const double2*)qa, (const double2*)qb, s2};
  return run_pairs(A, op, true, n, n, (double2*)gw, nullptr, nullptr);
}

extern "C" int eb_adj_stokes_pairs_vjp(eb_adj A, int32_t n, const double* z, const double* qa, const double* qb,
                                       double s2, const double* gwbar, double* zbar, double* qabar, double* qbbar) {
  if (!A || !z || !qa || !qb || !gwbar || !qabar || !qbbar)
    return (eb_set_error("eb_adj_stokes_pairs_vjp: bad arguments"), EB_ERR_ARG);
  int rc;
  // source side first (it writes zbar), then the target side into the scratch cf, added below
  StokesVjpS ops{(const double2*)z, (const double2*)qa, (const double2*)qb, (const double2*)gwbar, s2};
  if ((rc = run_pairs(A, ops, true, n, n, (double2*)zbar, (double2*)qabar, (double2*)qbbar)) != EB_OK) return rc;
  if (zbar) {
    if (n > A->n) return (eb_set_error("eb_adj_stokes_pairs_vjp: n above the handle's points"), EB_ERR_ARG);
    StokesVjpT opt{(const double2*)z, (const double2*)qa, (const double2*)qb, (const double2*)gwbar, s2};
    if ((rc = run_pairs(A, opt, true, n, n, A->cf, nullptr, nullptr)) != EB_OK) return rc;
    hipLaunchKernelGGL(k_add_to, dim3(ewb(n)), dim3(EWT), 0, A->st, n, (const double2*)A->cf, (double2*)zbar);
    EB_HIP_CHECK(hipGetLastError());
  }
  return EB_OK;
}
