// Host check: tuv_corner_range_of (per t-plane the least and greatest corner
// from the min / max of the two u terms and the two v terms) against the
// definition, the min and max over the eight corners evaluated one by one (as
// the oracle states it), bit for bit up to the sign of a zero, on 4M random
// pairs and boxes: unit scale, 1e-300 and 1e200 scales, coordinates on a
// coarse grid (ties and exact zeros), coincident vertices, signed zeros and
// the unit cube.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <random>
#include "cert_tuv_math.h"
using namespace p2c;
template <bool kEE>
void definition_range(const TuvPair& q, double tl, double th, double ul, double uh, double vl, double vh, double mn[3], double mx[3]) {
  const double ts[2] = {tl, th};
  for (int k = 0; k < 3; ++k) {
    double lo = INFINITY, hi = -INFINITY;
    for (int it = 0; it < 2; ++it) {
      const double X0 = q.x0[0][k] + ts[it] * q.d[0][k];
      const double X1 = q.x0[1][k] + ts[it] * q.d[1][k];
      const double X2 = q.x0[2][k] + ts[it] * q.d[2][k];
      const double X3 = q.x0[3][k] + ts[it] * q.d[3][k];
      double f[4];
      if (!kEE) {
        const double e1 = X2 - X1, e2 = X3 - X1;
        const double w1l = X1 + ul * e1, w1h = X1 + uh * e1;
        const double s2l = vl * e2, s2h = vh * e2;
        f[0] = X0 - (w1l + s2l); f[1] = X0 - (w1l + s2h); f[2] = X0 - (w1h + s2l); f[3] = X0 - (w1h + s2h);
      } else {
        const double ea = X1 - X0, eb = X3 - X2;
        const double pal = X0 + ul * ea, pah = X0 + uh * ea;
        const double pbl = X2 + vl * eb, pbh = X2 + vh * eb;
        f[0] = pal - pbl; f[1] = pal - pbh; f[2] = pah - pbl; f[3] = pah - pbh;
      }
      for (int c = 0; c < 4; ++c) { lo = fmin(lo, f[c]); hi = fmax(hi, f[c]); }
    }
    mn[k] = lo; mx[k] = hi;
  }
}
static bool same(double a, double b) { return a == b || (std::isnan(a) && std::isnan(b)); }
int main() {
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  long bad = 0, n = 0;
  for (long it = 0; it < 4000000; ++it) {
    TuvPair q; memset(&q, 0, sizeof q);
    const int mode = it % 8;
    const double scale = mode == 0 ? 1e-300 : mode == 1 ? 1e200 : mode == 2 ? 1.0 : mode == 3 ? 1e-8 : 1.0;
    for (int i = 0; i < 4; ++i) for (int k = 0; k < 3; ++k) {
      double a = U(rng) * scale, b = U(rng) * scale;
      if (mode == 4) { a = std::round(a * 4) / 4; b = std::round(b * 4) / 4; }       // ties and exact zeros
      if (mode == 5 && i == 1) a = q.x0[0][k];                                        // coincident vertices
      if (mode == 6) { a = (i & 1) ? 0.0 : -0.0; b = 0.0; }                           // signed zeros
      q.x0[i][k] = a; q.d[i][k] = b - a;
    }
    // a dyadic box at random depth
    int e[3]; uint64_t key[3];
    for (int d = 0; d < 3; ++d) { e[d] = rng() % 53; key[d] = (rng() >> 1) & ~((kOne63 >> e[d]) - 1) & (kOne63 - 1); if (e[d] == 0) key[d] = 0; }
    if (mode == 7) { e[0] = e[1] = e[2] = 0; key[0] = key[1] = key[2] = 0; }
    const double tl = key_to_f64(key[0]), ul = key_to_f64(key[1]), vl = key_to_f64(key[2]);
    const double th = tl + ldexp(1.0, -e[0]), uh = ul + ldexp(1.0, -e[1]), vh = vl + ldexp(1.0, -e[2]);
    for (int ee = 0; ee < 2; ++ee) {
      double a[3], b[3], c[3], d[3];
      if (ee) { definition_range<true>(q, tl, th, ul, uh, vl, vh, a, b); tuv_corner_range_of<true>(q, tl, th, ul, uh, vl, vh, c, d); }
      else { definition_range<false>(q, tl, th, ul, uh, vl, vh, a, b); tuv_corner_range_of<false>(q, tl, th, ul, uh, vl, vh, c, d); }
      ++n;
      for (int k = 0; k < 3; ++k) if (!same(a[k], c[k]) || !same(b[k], d[k])) {
        if (++bad < 5) printf("mode %d ee %d k %d: old [%.17g %.17g] new [%.17g %.17g]\n", mode, ee, k, a[k], b[k], c[k], d[k]);
      }
    }
  }
  printf("%ld evaluations, %ld mismatches\n", n, bad);
  return bad != 0;
}
