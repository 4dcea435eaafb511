// Sequential host emulation of one spectral workgroup running the per-body FFT
// (electricbubbles_amd/csrc/eb_fft.h) with the plans of eb_fft_plan.h.  Every stage runs
// its load phase for all threads, then its store phase, as the barrier between them orders
// them on the GPU.  Test infrastructure only: built by tests/test_fft_plan_cpu.py.
#include <string.h>

#include <vector>

#include "eb_fft_plan.h"

using namespace ebfft;

template <int R>
static void emu_spec(double2* x, int L, int nseq, const Stage& s, const double2* tab, int sign) {
  const int nbf = nseq * s.LR;
  std::vector<double2> v((size_t)nbf * R);
  for (int t = 0; t < nbf; ++t) spec_load<R>(x, L, t, s, tab + s.tw, sign, &v[(size_t)t * R]);
  for (int t = 0; t < nbf; ++t) {
    bfly<R>(&v[(size_t)t * R], (double)sign);
    spec_store<R>(x, L, t, s, &v[(size_t)t * R]);
  }
}

static void emu_stages(double2* x, int L, int nseq, const Tables& T, int st, int nst, int root, int sign) {
  const double2* tab = T.tab.data();
  for (int i = 0; i < nst; ++i) {
    const Stage& s = T.stages[(size_t)(st + i)];
    switch (s.R) {
      case 2: emu_spec<2>(x, L, nseq, s, tab, sign); break;
      case 3: emu_spec<3>(x, L, nseq, s, tab, sign); break;
      case 4: emu_spec<4>(x, L, nseq, s, tab, sign); break;
      case 5: emu_spec<5>(x, L, nseq, s, tab, sign); break;
      case 8: emu_spec<8>(x, L, nseq, s, tab, sign); break;
      default: {
        std::vector<double2> y((size_t)L);
        for (int o = 0; o < L; ++o) y[(size_t)o] = gen_out(x, L, o, s, tab + root, sign);
        memcpy(x, y.data(), sizeof(double2) * (size_t)L);
      }
    }
  }
}

extern "C" {

// Transforms n sequences of the given lengths (back to back, interleaved re/im) in place.
// info[0..3] of the last length: Bluestein P, Ns, M and the number of stages.
int fft_emul(int nbody, const int* lens, double* data, int sign, int ext_budget, int* info) {
  Tables T;
  size_t off = 0;
  for (int b = 0; b < nbody; ++b) {
    const int L = lens[b];
    const Plan p = T.plans[(size_t)plan_for(T, L, ext_budget)];
    double2* x = (double2*)data + off;
    emu_stages(x, L, 1, T, p.st, p.nst, p.root, sign);
    if (p.P) {
      const int E = p.Ns * p.M;
      std::vector<double2> U((size_t)E);
      for (int e = 0; e < E; ++e) U[(size_t)e] = blue_in(x, e, p, T.tab.data(), sign);
      emu_stages(U.data(), p.M, p.Ns, T, p.mst, p.mnst, 0, -1);
      for (int e = 0; e < E; ++e) U[(size_t)e] = blue_mul(U.data(), e, p, T.tab.data());
      emu_stages(U.data(), p.M, p.Ns, T, p.mst, p.mnst, 0, +1);
      for (int o = 0; o < L; ++o) x[o] = blue_out(U.data(), o, p, T.tab.data(), sign);
    }
    info[0] = p.P;
    info[1] = p.Ns;
    info[2] = p.M;
    info[3] = p.nst;
    off += (size_t)L;
  }
  return 0;
}

// t / d by the magic multiplier, for the exhaustive check of the division
int fdiv_emul(int t, int d) { return fdivm(t, d, magic_for(d)); }
}
