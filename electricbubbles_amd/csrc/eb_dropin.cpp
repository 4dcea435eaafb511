// eb_dropin.cpp -- host-buffer drop-in entry points with the reference's CADAPDR interface
// (cadapdrop.f:3-5; CADAPMOD cadapmod.f:3-5; CADAP cadap.f:3-5).
//
// The call copies the inputs to the GPU, evaluates GRADW with the device plan of
// eb_fmm.hip, copies GRADW back and returns IER(1).  Geometry is fixed across the matvecs of
// one GMRES solve (MATVEC re-sends the same XAT/YAT/DZDT every iteration,
// stokesbubb.f:449-455), so the plan -- the binned tree and the quadrature weights -- is
// kept and reused while (XAT, YAT, DZDT, ND, H, LEVMAX, TOL) are unchanged; then only DENS
// travels to the device and GRADW back.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "../../include/eb_atree.h"
#include "../../include/eb_cadap.h"

extern "C" void eb_set_error(const char* fmt, ...);
extern "C" int eb_fmm_plan_ntbox(eb_fmm_plan P, int32_t* ntbox);

namespace {

constexpr int kDirectMax = EB_DIRECT_MAX;  // EB_MODE_AUTO: O(N^2) sum at or below this many points

// One cached plan per variant: within one F the reference calls CADAPDR (SOLVEI's matvecs)
// and then CADAPMOD (INTERFACENEW, electric.f:835-837) on the same geometry, and a single slot
// would rebuild both plans on every F.
struct Slot {
  eb_fmm_plan plan = nullptr;
  int n = 0, levmax = 0, nterms = 0, pmode = -1;
  std::vector<double> kx, ky, kw;
  std::vector<int32_t> ntbox;
};

// DAPIF2's adaptive tree of the last geometry, for its INFORM and workspace figures
struct TreeKey {
  std::vector<double> x, y;
  int napb = -1, lmx = -1;
  long long nboxmx = -1;
  double eps7 = 0.0;
  int rc = -1;               // eb_atree_build's status
  int32_t terr[3] = {0, 0, 0};
  int32_t counts[8] = {0};   // nbox, nlev, len, len3, len4, maxp, jatlen, nat
};

struct Ctx {
  std::mutex mu;
  int mode = EB_MODE_AUTO;
  Slot slot[4];  // indexed by EB_VARIANT_*
  TreeKey tree;
  // device buffers shared by the slots (a plan reads x, y, wdz only while it bins them)
  int cap = 0;
  double *dx = nullptr, *dy = nullptr, *dw = nullptr, *ddens = nullptr, *dgw = nullptr;
  hipStream_t st = nullptr;
};

Ctx& ctx() {
  static Ctx* c = new Ctx();  // never destroyed: HIP may be torn down before static dtors run
  return *c;
}

int grow(Ctx& c, int n) {
  if (n <= c.cap) return EB_OK;
  if (c.st) (void)hipStreamSynchronize(c.st);
  double** bufs[] = {&c.dx, &c.dy, &c.dw, &c.ddens, &c.dgw};
  const size_t sz[] = {1, 1, 2, 2, 2};
  for (int i = 0; i < 5; ++i) {
    if (*bufs[i]) (void)hipFree(*bufs[i]);
    *bufs[i] = nullptr;
    if (hipMalloc((void**)bufs[i], sizeof(double) * sz[i] * (size_t)n) != hipSuccess) {
      eb_set_error("drop-in: hipMalloc of %d points failed", n);
      c.cap = 0;
      return EB_ERR_NOMEM;
    }
  }
  c.cap = n;
  return EB_OK;
}

int ensure_stream(Ctx& c) {
  if (c.st) return EB_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    eb_set_error("drop-in: no HIP device available");
    return EB_ERR_HIP;
  }
  if (hipStreamCreate(&c.st) != hipSuccess) {
    eb_set_error("drop-in: hipStreamCreate failed");
    return EB_ERR_HIP;
  }
  return EB_OK;
}

// Make slot `variant` hold a plan of n points at (x, y, w) (host arrays; w = weighted tangent,
// interleaved complex) with LEVMAX L, nterms terms and plan mode pmode.  A plan whose sizes
// match is re-binned in place (eb_fmm_plan_rebuild: no allocation); otherwise it is replaced.
// Called with the context lock held.
int slot_plan(Ctx& c, int variant, int n, const double* x, const double* y, const double* w, int L, int nterms,
              int pmode) {
  Slot& S = c.slot[variant];
  const bool sizes = S.plan && S.n == n && S.levmax == L && S.nterms == nterms && S.pmode == pmode;
  if (sizes && memcmp(S.kx.data(), x, sizeof(double) * n) == 0 && memcmp(S.ky.data(), y, sizeof(double) * n) == 0 &&
      memcmp(S.kw.data(), w, sizeof(double) * 2 * n) == 0)
    return EB_OK;
  int rc;
  if ((rc = grow(c, n)) != EB_OK) return rc;
  if (hipMemcpyAsync(c.dx, x, sizeof(double) * n, hipMemcpyHostToDevice, c.st) != hipSuccess ||
      hipMemcpyAsync(c.dy, y, sizeof(double) * n, hipMemcpyHostToDevice, c.st) != hipSuccess ||
      hipMemcpyAsync(c.dw, w, sizeof(double) * 2 * n, hipMemcpyHostToDevice, c.st) != hipSuccess) {
    eb_set_error("drop-in: host-to-device copy failed");
    return EB_ERR_HIP;
  }
  if (sizes) {
    rc = eb_fmm_plan_rebuild(S.plan, c.dx, c.dy, c.dw, c.st);
  } else {
    if (S.plan) eb_fmm_plan_destroy(S.plan);
    S.plan = nullptr;
    rc = eb_fmm_plan_create(&S.plan, n, c.dx, c.dy, c.dw, L, nterms, variant, pmode, c.st);
    if (rc != EB_OK) S.plan = nullptr;
  }
  if (rc != EB_OK) {
    if (S.plan) eb_fmm_plan_destroy(S.plan);
    S.plan = nullptr;
    return rc;
  }
  S.ntbox.assign(L + 1, 0);
  if ((rc = eb_fmm_plan_ntbox(S.plan, S.ntbox.data())) != EB_OK) {
    eb_fmm_plan_destroy(S.plan);
    S.plan = nullptr;
    return rc;
  }
  S.n = n;
  S.levmax = L;
  S.nterms = nterms;
  S.pmode = pmode;
  S.kx.assign(x, x + n);
  S.ky.assign(y, y + n);
  S.kw.assign(w, w + 2 * (size_t)n);
  return EB_OK;
}

// GW of slot `variant` for the density dens (host, interleaved complex) into gw (host).
int slot_apply(Ctx& c, int variant, int n, const double* dens, double* gw) {
  if (hipMemcpyAsync(c.ddens, dens, sizeof(double) * 2 * n, hipMemcpyHostToDevice, c.st) != hipSuccess) {
    eb_set_error("drop-in: density copy failed");
    return EB_ERR_HIP;
  }
  int rc = eb_fmm_apply(c.slot[variant].plan, c.ddens, c.dgw, c.st);
  if (rc != EB_OK) return rc;
  if (hipMemcpyAsync(gw, c.dgw, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, c.st) != hipSuccess ||
      hipStreamSynchronize(c.st) != hipSuccess) {
    eb_set_error("drop-in: device-to-host copy failed");
    return EB_ERR_HIP;
  }
  return EB_OK;
}

int32_t cadap_common(int variant, const int32_t* iflag1, const int32_t* iflag2, const int32_t* levmax,
                     const double* xat, const double* yat, const double* dzdt, const int32_t* k0,
                     const int32_t* kbod, const int32_t* nd, const double* dens, const double* h,
                     const int32_t* leniw, const int32_t* lenw, double* gradw, const double* tol,
                     int32_t* inform, int32_t* ier) {
  for (int i = 0; i < 5; ++i) ier[i] = 0;
  if (*iflag1 != 0 || *iflag2 != 0) {
    eb_set_error("CADAP drop-in: IFLAG1=%d IFLAG2=%d not supported (only 0, 0)", *iflag1, *iflag2);
    ier[0] = EB_ERR_UNSUPPORTED;
    return EB_ERR_UNSUPPORTED;
  }
  // NTERMS = -NINT(log(TOL)/log(2)) (cadapdrop.f:126-133)
  const int nterms = -(int)lround(log(*tol) / log(2.0));
  inform[1] = nterms;
  if (nterms > 60) {
    ier[0] = 4;
    ier[1] = nterms;
    return 4;
  }
  const int L = *levmax;
  // ND(K0:*) / H(K0:*) point at element K0
  const int nbod = *kbod - *k0 + 1;
  long long natoms = 0;
  for (int k = 0; k < nbod; ++k) natoms += nd[k];
  if (nbod < 1 || natoms < 2 || natoms > (1LL << 30) || L < 1 || L > 10 || nterms < 1) {
    eb_set_error("CADAP drop-in: unsupported sizes (bodies=%d points=%lld levmax=%d nterms=%d)", nbod, natoms, L,
                 nterms);
    ier[0] = EB_ERR_ARG;
    return EB_ERR_ARG;
  }
  const int n = (int)natoms;
  // complex workspace: ZSRC, QA, QB, QC and their gathers + ZTARG (cadapdrop.f:178-197);
  // NTARG is not read when IFLAG1 = 0 and counts as 0.
  const long long itotp = 8LL * n + 1;
  if (itotp > *lenw) {
    ier[0] = 8;
    ier[1] = (int32_t)itotp;
    ier[2] = *lenw;
    return 8;
  }

  Ctx& c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  int rc;
  if ((rc = ensure_stream(c)) != EB_OK) {
    ier[0] = rc;
    return rc;
  }
  const int pmode = (c.mode == EB_MODE_AUTO) ? (n <= kDirectMax ? EB_MODE_DIRECT : EB_MODE_FMM) : c.mode;

  // quadrature-weighted tangent DZDT * H(body), interleaved complex
  std::vector<double> w(2 * (size_t)n);
  {
    int j = 0;
    for (int k = 0; k < nbod; ++k)
      for (int i = 0; i < nd[k]; ++i, ++j) {
        w[2 * j] = dzdt[2 * j] * h[k];
        w[2 * j + 1] = dzdt[2 * j + 1] * h[k];
      }
  }
  if ((rc = slot_plan(c, variant, n, xat, yat, w.data(), L, nterms, pmode)) != EB_OK) {
    ier[0] = rc;
    return rc;
  }
  const Slot& S = c.slot[variant];
  // per-level workspace checks in the reference's order (cadapdrop.f:261-319): integer
  // tables first, then the local-expansion storage 2 (NTERMS+1) NTBOX.
  for (int l = 1; l <= L; ++l) {
    inform[0] = l;
    const long long nb = 1LL << (2 * l);
    const long long itot = 5 * nb + 2LL * n + 3;
    if (itot > *leniw) {
      ier[0] = 8;
      ier[3] = (int32_t)itot;
      ier[4] = *leniw;
      return 8;
    }
    const long long itotc = itotp + 2LL * (nterms + 1) * S.ntbox[l];
    if (itotc > *lenw) {
      ier[0] = 8;
      ier[1] = (int32_t)itotc;
      ier[2] = *lenw;
      return 8;
    }
  }
  if ((rc = slot_apply(c, variant, n, dens, gradw)) != EB_OK) {
    ier[0] = rc;
    return rc;
  }
  inform[0] = L;
  return EB_OK;
}

}  // namespace

#define EB_DROPIN(NAME, VARIANT)                                                                                  \
  extern "C" int32_t NAME(const int32_t* iflag1, const int32_t* iflag2, const int32_t* levmax, const double* xat,  \
                          const double* yat, const double* dzdt, const int32_t* k0, const int32_t* kbod,          \
                          const int32_t* nd, const double* dens, const double* h, int32_t* iwork,                 \
                          const int32_t* leniw, double* cwork, const int32_t* lenw, const double* xtarg,          \
                          const double* ytarg, const int32_t* ntarg, double* gradw, double* phip, double* phipp,  \
                          const double* tol, int32_t* inform, int32_t* ier, const double* xleft,                  \
                          const double* slength) {                                                                \
    (void)iwork; (void)cwork; (void)xtarg; (void)ytarg; (void)ntarg; (void)phip; (void)phipp;                    \
    (void)xleft; (void)slength;                                                                                  \
    return cadap_common(VARIANT, iflag1, iflag2, levmax, xat, yat, dzdt, k0, kbod, nd, dens, h, leniw, lenw,     \
                        gradw, tol, inform, ier);                                                                \
  }

EB_DROPIN(eb_cadapdr, EB_VARIANT_DR)
EB_DROPIN(eb_cadapmod, EB_VARIANT_MOD)
EB_DROPIN(eb_cadap, EB_VARIANT_PLUS)

// ------------------------------------------------------------------------------------------
// DAPIF2 (dapif1.f:458-999) with IFLAG7 = 3, the Hilbert-matrix product
//   FIELD(i) = sum_{j != i} QA(j) / (z_i - z_j)
// that FASMVP_LAP (laplace.f:204-215) and PVINTEV (pv_integral.f:171-182) call.  FIELD comes from
// the GPU FMM (EB_VARIANT_HILBERT with unit weights, LEVMAX 7 and -NINT(log2 TOL) terms, as F's
// own Laplace products), or from the O(N^2) sum as the drop-in mode says.  IERR and INFORM are
// the reference's: NTERMS = -NINT(log2 TOL) + 3 (IERR 4 above 60), IERR 16 when NAT <= NAPB,
// the adaptive tree of DAPIF2's rules built on the GPU (eb_atree: NBOX, the level count and
// IERR 1 / 32 bit for bit), and the workspace layout of dapif1.f:745-958 address by address,
// with IERR 8 and IERR(2) = the address needed at the first allocation that does not fit,
// DAFAMI's own test of the tree's share before compression (dapif1.f:1436-1441) included, and
// INFORM(4) = the workspace used.
// IFLAG7 = 1 or 2 (fields and potentials of real charges) have no caller on the path and
// return EB_ERR_UNSUPPORTED with IERR(1) = -1.
// ------------------------------------------------------------------------------------------
namespace {

int tree_for(Ctx& c, int nat, const double* xa, const double* ya, int napb, int lmx, double eps7, long long nboxmx) {
  TreeKey& T = c.tree;
  if (T.rc >= 0 && (int)T.x.size() == nat && T.napb == napb && T.lmx == lmx && T.eps7 == eps7 &&
      T.nboxmx == nboxmx &&
      memcmp(T.x.data(), xa, sizeof(double) * nat) == 0 && memcmp(T.y.data(), ya, sizeof(double) * nat) == 0)
    return EB_OK;
  int rc;
  if ((rc = grow(c, nat)) != EB_OK) return rc;
  if (hipMemcpyAsync(c.dx, xa, sizeof(double) * nat, hipMemcpyHostToDevice, c.st) != hipSuccess ||
      hipMemcpyAsync(c.dy, ya, sizeof(double) * nat, hipMemcpyHostToDevice, c.st) != hipSuccess) {
    eb_set_error("DAPIF2 drop-in: host-to-device copy failed");
    return EB_ERR_HIP;
  }
  eb_atree t = nullptr;
  int32_t terr[3] = {0, 0, 0};
  rc = eb_atree_build_bounded(&t, nat, c.dx, c.dy, napb, lmx, eps7, nboxmx, terr, c.st);
  T.rc = -1;
  if (rc != EB_OK && rc != EB_ATREE_EPS_STOP && rc != EB_ATREE_TOO_DEEP && rc != EB_ATREE_FEW_POINTS &&
      rc != EB_ATREE_WORKSPACE)
    return rc;
  if (t) {
    rc = eb_atree_counts(t, T.counts);
    eb_atree_destroy(t);
    if (rc != EB_OK) return rc;
    T.rc = EB_OK;
  } else {
    T.rc = rc;  // a tree error: IERR as DAPIF2 sets it
  }
  memcpy(T.terr, terr, sizeof(terr));
  T.x.assign(xa, xa + nat);
  T.y.assign(ya, ya + nat);
  T.napb = napb;
  T.lmx = lmx;
  T.eps7 = eps7;
  T.nboxmx = nboxmx;
  return EB_OK;
}

}  // namespace

extern "C" int32_t eb_dapif2(int32_t* iout, const int32_t* iflag7, const int32_t* nat, const int32_t* napb,
                             const int32_t* ninire, const int32_t* mex, int32_t* ierr, int32_t* inform,
                             const double* tol, const double* eps7, const double* xa, const double* ya,
                             const double* qa, double* poten, double* field, double* wksp, const int32_t* nsp,
                             const void* clos) {
  (void)iout;
  (void)poten;
  (void)wksp;
  (void)clos;  // CLOSE is never called (the callers pass an undeclared variable)
  if (*iflag7 != 3) {
    ierr[0] = -1;
    eb_set_error("DAPIF2 drop-in: IFLAG7 = %d not supported (only 3, the Hilbert-matrix product)", *iflag7);
    return EB_ERR_UNSUPPORTED;
  }
  const long long NAT = *nat, NAPB = *napb, NI = *ninire, NSP = *nsp;
  if (NAT < 2 || NAT > (1LL << 30) || NI < 1) {
    ierr[0] = -1;
    eb_set_error("DAPIF2 drop-in: unsupported sizes (NAT=%lld NINIRE=%lld)", NAT, NI);
    return EB_ERR_ARG;
  }
  // dapif1.f:680-714: terms, the level limit, NAT > NAPB
  ierr[0] = 0;
  ierr[2] = (int32_t)NSP;
  const int N = -(int)lround(log(*tol) / log(2.0)) + 3;
  inform[0] = N;
  if (N > 60) {
    ierr[0] = 4;
    ierr[1] = N;
    return 4;
  }
  const double rex2 = *mex * (log(10.0) / log(2.0)) - 1.0;
  const int lmx = (int)(rex2 / N + 6.0);
  if (NAT <= NAPB) {
    ierr[0] = 16;
    ierr[1] = (int32_t)NAT;
    ierr[2] = (int32_t)NAPB;
    return 16;
  }
  // the workspace, address by address (1-based REAL*8 words); IERR(2) = the address needed - 1
  long long adr = 1 + 2 * NAT;  // the scaled coordinates (dapif1.f:752-762)
  auto fits = [&](long long a) {
    if (a - 1 <= NSP) return true;
    ierr[0] = 8;
    ierr[1] = (int32_t)(a - 1);
    return false;
  };
  if (!fits(adr)) return 8;
  adr += (NAT + NI - 1) / NI;  // IBOX (dapif1.f:777-784)
  if (!fits(adr)) return 8;
  const long long xb = adr;
  const long long nboxmx = (NSP - adr + 1) * NI / (2 * NI + 4);  // the tree's share (dapif1.f:784)
  Ctx& c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  int rc;
  if ((rc = ensure_stream(c)) != EB_OK || (rc = tree_for(c, (int)NAT, xa, ya, (int)NAPB, lmx, *eps7, nboxmx)) != EB_OK) {
    ierr[0] = -1;
    return rc;
  }
  if (c.tree.rc != EB_OK) {  // DAASSI's IERR 1 (smallest box below EPS7), 32 (too many levels), 8 (NBOXMX)
    // the words each error sets: EPS IERR(1); DAFAMI IERR(1..2); too deep IERR(1..3)
    const int nset = c.tree.rc == EB_ATREE_TOO_DEEP ? 3 : (c.tree.rc == EB_ATREE_WORKSPACE ? 2 : 1);
    for (int i = 0; i < nset; ++i) ierr[i] = c.tree.terr[i];
    return ierr[0];
  }
  const long long nbox = c.tree.counts[0], len = c.tree.counts[2], len3 = c.tree.counts[3],
                  len4 = c.tree.counts[4], maxp = c.tree.counts[5];
  inform[1] = (int32_t)nbox;
  inform[2] = c.tree.counts[1];
  const long long nboxi = (nbox + NI - 1) / NI;
  adr = xb + 2 * nbox + 3 * nboxi + 3;  // XB, YB, NBAT, NPAR, NCHI moved down (dapif1.f:810-849)
  adr += nboxi + 1;                     // INDB (dapif1.f:853-862)
  if (!fits(adr)) return 8;
  adr += (nbox + 1 + NI - 1) / NI + (nbox + 1 + NAT + NI - 1) / NI;  // IAT, JAT (dapif1.f:870-880)
  if (!fits(adr)) return 8;
  const long long jl = adr + 2 * (nboxi + 1);  // L1, L2, then JL (dapif1.f:888-897)
  if (!fits(jl + 36 * (nboxi + 1))) return 8;
  const long long idesc = jl + (len + NI - 1) / NI + 2 * (nboxi + 1);  // L3, L4 (dapif1.f:910-915)
  adr = idesc + (len3 + NI - 1) / NI + (len4 + NI - 1) / NI;          // JL3, JL4 compressed (dapif1.f:930-942)
  if (!fits(adr)) return 8;
  // LO, TA, IZ, ICH, IBAK, N2 (dapif1.f:961-968)
  adr += 4 * (nbox + 1) * N + 4 * maxp + 2 * ((maxp + NI - 1) / NI);
  inform[3] = (int32_t)adr;
  if (!fits(adr)) return 8;
  // FIELD on the GPU: unit weights, QA as the density
  const int n = (int)NAT;
  const int L = 7, nterms = N - 3;
  const int pmode = (c.mode == EB_MODE_AUTO) ? (n <= kDirectMax ? EB_MODE_DIRECT : EB_MODE_FMM) : c.mode;
  std::vector<double> w(2 * (size_t)n, 0.0);
  for (int i = 0; i < n; ++i) w[2 * i] = 1.0;
  if ((rc = slot_plan(c, EB_VARIANT_HILBERT, n, xa, ya, w.data(), L, nterms, pmode)) != EB_OK ||
      (rc = slot_apply(c, EB_VARIANT_HILBERT, n, qa, field)) != EB_OK) {
    ierr[0] = -1;
    return rc;
  }
  return EB_OK;
}

extern "C" int eb_set_mode(int mode) {
  if (mode != EB_MODE_AUTO && mode != EB_MODE_FMM && mode != EB_MODE_DIRECT) {
    eb_set_error("eb_set_mode: unknown mode %d", mode);
    return EB_ERR_ARG;
  }
  Ctx& c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  c.mode = mode;
  return EB_OK;
}

extern "C" int eb_get_mode(void) {
  Ctx& c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  return c.mode;
}
