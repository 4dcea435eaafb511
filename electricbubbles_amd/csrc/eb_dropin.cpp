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

struct Ctx {
  std::mutex mu;
  int mode = EB_MODE_AUTO;
  Slot slot[4];  // indexed by EB_VARIANT_*
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
