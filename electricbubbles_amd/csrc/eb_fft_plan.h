// eb_fft_plan.h -- host-side plans and tables of the per-body FFT (eb_fft.h).
//
// For every distinct body length the builder appends to three arrays shared by all bodies:
// the stage list, the Plan record and a table of complex constants (stage twiddles, roots
// of unity for direct-sum stages, Bluestein chirps and filter transforms), all computed in
// long double and rounded once to double.
#pragma once
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <complex>
#include <map>
#include <vector>

#include "eb_fft.h"

namespace ebfft {

using cld = std::complex<long double>;
static const long double PI_L = 3.14159265358979323846264338327950288L;

// exp(-2 pi i num / den) with num reduced mod den first (exact argument)
inline cld root_ld(long long num, long long den) {
  num %= den;
  if (num < 0) num += den;
  const long double a = -2.0L * PI_L * (long double)num / (long double)den;
  return cld(cosl(a), sinl(a));
}
inline double2 to_d2(cld z) { return make_double2((double)z.real(), (double)z.imag()); }

inline bool smooth235(int n) {
  for (int p : {2, 3, 5})
    while (n % p == 0) n /= p;
  return n == 1;
}
inline int next_smooth235(int n) {
  while (!smooth235(n)) ++n;
  return n;
}

// prime factors in ascending order, with multiplicity
inline std::vector<int> prime_factors(int n) {
  std::vector<int> f;
  for (int p = 2; p * p <= n; ++p)
    while (n % p == 0) {
      f.push_back(p);
      n /= p;
    }
  if (n > 1) f.push_back(n);
  return f;
}

// forward DFT of length n = 2^a 3^b 5^c in long double (recursive decimation in time); w is
// the table exp(-2 pi i j / n0) of the top-level length n0, n divides n0
inline void dft_ld_rec(const cld* in, int stride, int n, cld* out, const std::vector<cld>& w, int n0) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  const int p = (n % 2 == 0) ? 2 : (n % 3 == 0) ? 3 : 5;
  const int m = n / p;
  std::vector<cld> sub((size_t)n);
  for (int r = 0; r < p; ++r) dft_ld_rec(in + (size_t)r * stride, stride * p, m, &sub[(size_t)r * m], w, n0);
  const int scale = n0 / n;
  for (int q = 0; q < m; ++q)
    for (int s = 0; s < p; ++s) {
      cld acc = 0;
      for (int r = 0; r < p; ++r) acc += sub[(size_t)r * m + q] * w[(size_t)(((long long)r * (q + s * m)) % n) * scale];
      out[q + s * m] = acc;
    }
}

struct Tables {
  std::vector<Stage> stages;
  std::vector<Plan> plans;
  std::vector<double2> tab;
  std::map<int, int> plan_of_len;  // L -> index into plans
};

// closed-form stages of a length with factors 2, 3, 5 (and direct-sum stages for the rest)
inline void add_stages(Tables& T, int L, const std::vector<int>& radices, int* st, int* nst, int* cost) {
  *st = (int)T.stages.size();
  *nst = (int)radices.size();
  int Ns = 1;
  for (int R : radices) {
    Stage s;
    s.R = R;
    s.Ns = Ns;
    s.LR = L / R;
    s.mNs = magic_for(Ns);
    s.mLR = magic_for(s.LR);
    s.tw = (int)T.tab.size();
    const bool closed = (R == 2 || R == 3 || R == 4 || R == 5 || R == 8);
    if (closed && Ns > 1) {
      for (int r = 1; r < R; ++r)
        for (int k = 0; k < Ns; ++k) T.tab.push_back(to_d2(root_ld((long long)k * r, (long long)Ns * R)));
    }
    *cost += closed ? L : L * R / 2;
    T.stages.push_back(s);
    Ns *= R;
  }
}

// radix list for the 2, 3, 5 part of n: 8s, then a 4 or a 2 for the rest of the twos, 3s, 5s
inline std::vector<int> radices235(int n2, int n3, int n5) {
  std::vector<int> r;
  for (int i = 0; i < n2 / 3; ++i) r.push_back(8);
  if (n2 % 3 == 2) r.push_back(4);
  if (n2 % 3 == 1) r.push_back(2);
  for (int i = 0; i < n3; ++i) r.push_back(3);
  for (int i = 0; i < n5; ++i) r.push_back(5);
  return r;
}

// Plan index for length L; ext_budget = LDS entries available for the Bluestein workspace
inline int plan_for(Tables& T, int L, int ext_budget) {
  auto it = T.plan_of_len.find(L);
  if (it != T.plan_of_len.end()) return it->second;
  Plan p{};
  p.L = L;
  std::vector<int> pf = prime_factors(L);
  int n2 = 0, n3 = 0, n5 = 0;
  std::vector<int> other;
  for (int f : pf) {
    if (f == 2) ++n2;
    else if (f == 3) ++n3;
    else if (f == 5) ++n5;
    else other.push_back(f);
  }
  // Bluestein for the largest prime, when it is large and its workspace fits
  int P = 0, Ns = 0, M = 0;
  if (!other.empty() && other.back() >= BL_MIN) {
    const int cand = other.back();
    const int m = next_smooth235(2 * cand - 1);
    const int ns = L / cand;
    if ((long long)ns * m <= std::min(EMAX, ext_budget)) {
      P = cand;
      Ns = ns;
      M = m;
      other.pop_back();
    }
  }
  std::vector<int> radices = radices235(n2, n3, n5);
  for (int f : other) radices.push_back(f);
  int cost = 0;
  add_stages(T, L, radices, &p.st, &p.nst, &cost);
  p.root = (int)T.tab.size();
  if (!other.empty())  // primes of 7 and up run as direct sums over the roots of unity
    for (int j = 0; j < L; ++j) T.tab.push_back(to_d2(root_ld(j, L)));
  if (P) {
    p.P = P;
    p.Ns = Ns;
    p.M = M;
    p.mM = magic_for(M);
    p.mNsB = magic_for(Ns);
    // c_n = exp(-i pi n^2 / P) = root(n^2 mod 2P, 2P)
    auto chirp = [&](long long n) { return root_ld((n * n) % (2LL * P), 2LL * P); };
    p.pre = (int)T.tab.size();
    for (int s = 0; s < Ns; ++s)
      for (int r = 0; r < P; ++r) {
        // W_L^{r s} c_r as one angle: -2 pi (2 P (r s mod L) + L (r^2 mod 2P)) / (2 P L)
        const long long a = 2LL * P * (((long long)r * s) % L) + (long long)L * (((long long)r * r) % (2LL * P));
        T.tab.push_back(to_d2(root_ld(a, 2LL * P * L)));
      }
    p.post = (int)T.tab.size();
    for (int q = 0; q < P; ++q) T.tab.push_back(to_d2(chirp(q)));
    // filter h_m = conj(c_m) for m < P and h_{M-m} = conj(c_m) for 0 < m < P; H = DFT_M(h) / M
    std::vector<cld> h((size_t)M, cld(0)), Hf((size_t)M), w((size_t)M);
    for (int m = 0; m < P; ++m) h[(size_t)m] = std::conj(chirp(m));
    for (int m = 1; m < P; ++m) h[(size_t)(M - m)] = std::conj(chirp(m));
    for (int j = 0; j < M; ++j) w[(size_t)j] = root_ld(j, M);
    dft_ld_rec(h.data(), 1, M, Hf.data(), w, M);
    p.H = (int)T.tab.size();
    for (int m = 0; m < M; ++m) T.tab.push_back(to_d2(Hf[(size_t)m] / (long double)M));
    int m2 = 0, m3 = 0, m5 = 0, mm = M;
    while (mm % 2 == 0) { ++m2; mm /= 2; }
    while (mm % 3 == 0) { ++m3; mm /= 3; }
    while (mm % 5 == 0) { ++m5; mm /= 5; }
    int mcost = 0;
    add_stages(T, M, radices235(m2, m3, m5), &p.mst, &p.mnst, &mcost);
    cost += Ns * (2 * mcost + 3 * M) + L;
  }
  p.cost = cost;
  const int idx = (int)T.plans.size();
  T.plans.push_back(p);
  T.plan_of_len[L] = idx;
  return idx;
}

}  // namespace ebfft
