// Host check: tuv_split_sequence (running ratios, one packed word per 32
// levels) against the level-by-level tuv_split_dim replay, on special values
// (0, 1, ties, powers of two, subnormals, inf, NaN) and 2M random inputs;
// then the one-hot level masks built from it (tuv_split_masks) against the
// packed words: each level's dimension and the exponents at each level.
#include <cstdio>
#include <cmath>
#include <random>
#include <cstring>
#include "cert_tuv_math.h"
using namespace p2c;
// The definitional replay: tuv_split_dim evaluated afresh at every level.
static void replay_seq(const double inv_tol[3], uint64_t out[5], uint32_t* dl) {
  TuvPair q; memset(&q, 0, sizeof q);
  for (int k = 0; k < 3; ++k) q.inv_tol[k] = inv_tol[k];
  uint64_t w0 = ~0ull, w1 = ~0ull, w2 = ~0ull, w3 = ~0ull, w4 = ~0ull;
  int e[3] = {0, 0, 0};
  uint32_t level = 0;
  for (;;) {
    const int d = tuv_split_dim(q, e);
    if (d == 3) break;
    const uint32_t wi = level >> 5, sh = (level & 31u) * 2u;
    const uint64_t clear = ~(3ull << sh), set = (uint64_t)d << sh;
    w0 = wi == 0 ? ((w0 & clear) | set) : w0;
    w1 = wi == 1 ? ((w1 & clear) | set) : w1;
    w2 = wi == 2 ? ((w2 & clear) | set) : w2;
    w3 = wi == 3 ? ((w3 & clear) | set) : w3;
    w4 = wi == 4 ? ((w4 & clear) | set) : w4;
    put3(e, d, get3(e, d) + 1);
    ++level;
  }
  out[0] = w0; out[1] = w1; out[2] = w2; out[3] = w3; out[4] = w4; *dl = level;
}
int main() {
  std::mt19937_64 rng(7);
  const double specials[] = {0.0, 1.0, 2.0, 0.5, 1.0000000000000002, 0.9999999999999999, 3.0, 4.0, 1e300, 1e-300,
                             4.9e-324, INFINITY, NAN, 1e6, 3e6, 2.5e15, 4503599627370496.0, 9007199254740992.0,
                             1e16, 1e20, 7.0, 8.0, 1.5, 6e5};
  const int ns = sizeof(specials) / sizeof(double);
  long bad = 0, n = 0;
  auto check = [&](double a, double b, double c) {
    double it[3] = {a, b, c};
    uint64_t o[5], w[5]; uint32_t dlo, dln;
    replay_seq(it, o, &dlo);
    tuv_split_sequence(it, w, &dln);
    ++n;
    if (dlo != dln || memcmp(o, w, sizeof o)) {
      if (bad < 5) printf("mismatch %g %g %g: %u vs %u\n", a, b, c, dlo, dln);
      ++bad;
      return;
    }
    // the level masks the kernels read: every level's dimension and the
    // exponents at every level up to done_level, against the 2-bit words
    TuvPair q; memset(&q, 0, sizeof q);
    tuv_split_masks(w, q.tmask, q.umask, &q.split_hi);
    q.done_level = dln;
    // the setup's direct builder: the same masks, high codes and length
    {
      uint64_t tm[2], um[2], sh;
      uint32_t dl;
      tuv_split_masks_of(it, tm, um, &sh, &dl);
      if (dl != dln || tm[0] != q.tmask[0] || tm[1] != q.tmask[1] || um[0] != q.umask[0] || um[1] != q.umask[1] ||
          sh != q.split_hi) {
        if (bad < 5) printf("direct mask mismatch %g %g %g\n", a, b, c);
        ++bad;
        return;
      }
    }
    int e[3] = {0, 0, 0};
    for (uint32_t l = 0; l <= dln; ++l) {
      int ex[3];
      tuv_exponents(q, l, ex);
      if (ex[0] != e[0] || ex[1] != e[1] || ex[2] != e[2]) {
        if (bad < 5) printf("exponent mismatch %g %g %g at level %u\n", a, b, c, l);
        ++bad;
        return;
      }
      if (l == dln) break;
      const int code = (int)((w[l >> 5] >> ((l & 31u) * 2u)) & 3ull);
      if (tuv_level_dim(q, l) != code) {
        if (bad < 5) printf("level mask mismatch %g %g %g at level %u\n", a, b, c, l);
        ++bad;
        return;
      }
      ++e[code];
    }
  };
  for (int i = 0; i < ns; ++i) for (int j = 0; j < ns; ++j) for (int k = 0; k < ns; ++k) check(specials[i], specials[j], specials[k]);
  std::uniform_real_distribution<double> U(-5, 25);
  for (int t = 0; t < 2000000; ++t) {
    double v[3];
    for (int k = 0; k < 3; ++k) {
      const int mode = rng() % 4;
      if (mode == 0) v[k] = std::pow(10.0, U(rng));
      else if (mode == 1) v[k] = std::ldexp(1.0, (int)(rng() % 80) - 10);        // exact powers of two
      else if (mode == 2) v[k] = std::ldexp(1.0 + (rng() % 8) / 8.0, (int)(rng() % 60)); // frequent ties
      else v[k] = (rng() % 2) ? v[0] : std::pow(10.0, U(rng));
    }
    check(v[0], v[1], v[2]);
  }
  printf("split_sequence_check: %ld cases, %ld mismatches\n", n, bad);
  return bad != 0;
}
