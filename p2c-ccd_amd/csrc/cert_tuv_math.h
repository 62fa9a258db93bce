// (t,u,v) inclusion certificate: per-pair setup, box evaluation and the
// stackless depth-first walk, shared by the gfx950 kernels.
//
// Method: the inclusion-based root finder of Wang et al., ACM TOG 2021 (the
// library the reference calls, src/tools/tight_inclusion_full_query_benchmark
// .cpp:375-397).  The certificate's definition -- multilinear corner
// enclosure, forward error bound, tolerance-driven split order, "toi = least
// t_lo over done boxes" -- is stated in full in the library's design notes
// and implemented here operation for operation (FP64, no contraction).
//
// Box encoding: lower corners as dyadic keys scaled by 2^63 (kt, ku, kv) and
// a level.  Every box at a level has the same widths, because the split
// dimension is a function of the widths alone; the per-pair split sequence is
// precomputed (one-hot level masks), so a box is (keys, level) and the path
// to it can be read back from the key bits.  The stream kernel's jump walk
// keeps 32-bit keys (scaled by 2^31) for the pairs whose dimensions split at
// most 31 times each.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define P2C_TUV_HD __host__ __device__ inline

namespace p2c {

constexpr double kTuvErrScale = 5.6e-15;   // >= 50 * 2^-53: first-order corner error per unit |x|
constexpr double kTuvErrFloor = 1e-300;    // covers underflow in the corner evaluation
constexpr int kTuvMaxExp = 52;             // per-dimension depth limit (keys stay exact in FP64)
constexpr int kTuvMaxLevel = 3 * kTuvMaxExp;
constexpr uint64_t kOne63 = 1ull << 63;
constexpr uint64_t kTuvNoHit = ~0ull;

constexpr uint32_t kTuvMaskLevels = 128;  // levels held as one-hot masks

// Per-pair constants (288 bytes).  The split sequence (the dimension each
// level splits) is held as one-hot level masks for levels 0..127 -- bit l of
// tmask (umask) set when level l splits t (u); a level below done_level with
// neither bit splits v -- and as 2-bit codes (3 = none) for levels 128..155,
// which only pairs with extreme extents reach.  Masks make the exponents of a
// level two population counts.
struct TuvPair {
  double x0[4][3];   // vertex positions at t = 0
  double d[4][3];    // x1 - x0
  double err[3];     // corner-evaluation error bound + minimum separation
  double inv_tol[3]; // 3 L_k / tolerance
  uint64_t tmask[2]; // levels 0..127 that split t
  uint64_t umask[2]; // levels 0..127 that split u
  uint64_t split_hi; // levels 128..155: split dimension, 2 bits each
  uint32_t done_level;
  uint32_t is_ee;
};

// k * 2^-63 for a box corner key: a multiple of 2^11 below 2^63 (depth <= 52
// per dimension), so k >> 11 < 2^52 sits exactly in the mantissa of 2^52
// and the conversion is exact (the same value as (double)k * 0x1p-63).
P2C_TUV_HD double key_to_f64(uint64_t k) {
  return (__builtin_bit_cast(double, 0x4330000000000000ull | (k >> 11)) - 0x1p52) * 0x1p-52;
}

// F at the 8 corners of box [tl,th]x[ul,uh]x[vl,vh]; per-coordinate min/max.
// The pair kind is a template parameter so the six (coordinate, time)
// blocks are straight-line code the compiler can interleave; a run-time
// branch inside each block would serialise them.
//
// On one t-plane the four corners are VF: X0 - (w1(u) + s2(v)), EE:
// pa(u) - pb(v), each the correctly rounded result of one operation on its
// two operands.  Round-to-nearest is monotone in each operand, so the least
// corner is X0 - (max w1 + max s2) (EE: min pa - max pb) and the greatest
// X0 - (min w1 + min s2) (EE: max pa - min pb): the same doubles as taking
// the min and max over the four corners evaluated one by one (the
// definition, as the oracle states it), up to the sign of a zero, for
// corner values that do not overflow.  Per plane that is 4 operations and
// 4 min/max instead of 8 operations and 8 min/max.
template <bool kEE>
P2C_TUV_HD void tuv_corner_range_of(const TuvPair& q, double tl, double th, double ul, double uh, double vl,
                                    double vh, double mn[3], double mx[3]) {
  const double ts[2] = {tl, th};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double lo[2], hi[2];
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      const double X0 = q.x0[0][k] + ts[it] * q.d[0][k];
      const double X1 = q.x0[1][k] + ts[it] * q.d[1][k];
      const double X2 = q.x0[2][k] + ts[it] * q.d[2][k];
      const double X3 = q.x0[3][k] + ts[it] * q.d[3][k];
      if (!kEE) {
        const double e1 = X2 - X1, e2 = X3 - X1;
        const double w1l = X1 + ul * e1, w1h = X1 + uh * e1;
        const double s2l = vl * e2, s2h = vh * e2;
        lo[it] = X0 - (__builtin_fmax(w1l, w1h) + __builtin_fmax(s2l, s2h));
        hi[it] = X0 - (__builtin_fmin(w1l, w1h) + __builtin_fmin(s2l, s2h));
      } else {
        const double ea = X1 - X0, eb = X3 - X2;
        const double pal = X0 + ul * ea, pah = X0 + uh * ea;
        const double pbl = X2 + vl * eb, pbh = X2 + vh * eb;
        lo[it] = __builtin_fmin(pal, pah) - __builtin_fmax(pbl, pbh);
        hi[it] = __builtin_fmax(pal, pah) - __builtin_fmin(pbl, pbh);
      }
    }
    mn[k] = __builtin_fmin(lo[0], lo[1]);
    mx[k] = __builtin_fmax(hi[0], hi[1]);
  }
}

P2C_TUV_HD void tuv_corner_range(const TuvPair& q, double tl, double th, double ul, double uh, double vl,
                                 double vh, double mn[3], double mx[3]) {
  if (q.is_ee)
    tuv_corner_range_of<true>(q, tl, th, ul, uh, vl, vh, mn, mx);
  else
    tuv_corner_range_of<false>(q, tl, th, ul, uh, vl, vh, mn, mx);
}

// Which dimension a box with exponents e splits along, or 3 when none may.
P2C_TUV_HD int tuv_split_dim(const TuvPair& q, const int e[3]) {
  int best = 3;
  double br = 1.0;
  for (int k = 0; k < 3; ++k) {
    if (e[k] >= kTuvMaxExp) continue;
    const double r = __builtin_ldexp(1.0, -e[k]) * q.inv_tol[k];
    if (r > br) {
      br = r;
      best = k;
    }
  }
  return best;
}

// Element d of a 3-vector through constant indices only: a dynamic index
// would move the array out of registers (to scratch, or to LDS when the
// compiler promotes it).
template <typename T>
P2C_TUV_HD T get3(const T v[3], int d) {
  return d == 0 ? v[0] : (d == 1 ? v[1] : v[2]);
}
template <typename T>
P2C_TUV_HD void put3(T v[3], int d, T x) {
  v[0] = d == 0 ? x : v[0];
  v[1] = d == 1 ? x : v[1];
  v[2] = d == 2 ? x : v[2];
}

// Split dimension of a level below kTuvMaskLevels (and below done_level).
// Words are chosen with constant indices so the pair record stays in
// registers (a dynamic index would move it to scratch memory).
P2C_TUV_HD int tuv_mask_dim(const TuvPair& q, uint32_t level) {
  const bool hi = level >= 64;
  const uint64_t bit = 1ull << (level & 63u);
  const uint64_t t = hi ? q.tmask[1] : q.tmask[0];
  const uint64_t u = hi ? q.umask[1] : q.umask[0];
  return (t & bit) ? 0 : ((u & bit) ? 1 : 2);
}

// Split dimension of `level` (below done_level); levels past the masks, which
// few walks reach, take a branch the wave skips when no lane needs it.
P2C_TUV_HD int tuv_level_dim(const TuvPair& q, uint32_t level) {
  if (level < kTuvMaskLevels) return tuv_mask_dim(q, level);
  return (int)((q.split_hi >> ((level - kTuvMaskLevels) * 2u)) & 3ull);
}

// Bits 0, 2, 4, ... of x (its odd bits zero) packed into bits 0..31.
P2C_TUV_HD uint64_t pack_even_bits(uint64_t x) {
  x = (x | (x >> 1)) & 0x3333333333333333ull;
  x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
  x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
  x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
  return x;
}

// The pair's level masks and high codes from the 2-bit split words (32
// levels per word, code 0 = t, 1 = u, 2 = v, 3 = none).
P2C_TUV_HD void tuv_split_masks(const uint64_t split[5], uint64_t tmask[2], uint64_t umask[2], uint64_t* split_hi) {
  constexpr uint64_t kEven = 0x5555555555555555ull;
  uint64_t t[4], u[4];
  for (int w = 0; w < 4; ++w) {
    const uint64_t b0 = split[w] & kEven, b1 = (split[w] >> 1) & kEven;
    t[w] = pack_even_bits(kEven & ~(b0 | b1));  // code 0
    u[w] = pack_even_bits(b0 & ~b1);            // code 1
  }
  tmask[0] = t[0] | (t[1] << 32);
  tmask[1] = t[2] | (t[3] << 32);
  umask[0] = u[0] | (u[1] << 32);
  umask[1] = u[2] | (u[3] << 32);
  *split_hi = split[4];
}

// Appends dimension d at `level` to the packed sequence (one register word,
// placed into w0..w4 when it fills).
P2C_TUV_HD void split_append(uint64_t& cur, uint32_t& level, int d, uint64_t& w0, uint64_t& w1, uint64_t& w2,
                             uint64_t& w3, uint64_t& w4) {
  const uint32_t sh = (level & 31u) * 2u;
  cur = (cur & ~(3ull << sh)) | ((uint64_t)d << sh);
  ++level;
  if ((level & 31u) == 0) {
    const uint32_t wi = (level >> 5) - 1;
    w0 = wi == 0 ? cur : w0;
    w1 = wi == 1 ? cur : w1;
    w2 = wi == 2 ? cur : w2;
    w3 = wi == 3 ? cur : w3;
    w4 = wi == 4 ? cur : w4;
    cur = ~0ull;
  }
}

// The split sequence of tuv_split_dim, 2 bits per level (3 = no split),
// and its length.  Each level splits the dimension with the largest
// r_k = 2^-e_k * inv_tol_k above 1 (e_k < kTuvMaxExp; the first such k on
// ties).  Writing inv_tol_k = m_k 2^E_k (m_k in [1, 2)), r_k > 1 holds for
// e_k < E_k and, when m_k > 1, for e_k = E_k: dimension k splits
// n_k = min(52, E_k + [m_k > 1]) times, its j-th split ranking by
// (E_k - j, m_k, -k).  So the sequence runs in rounds v = E_max, E_max - 1,
// ...: in round v every dimension with a split of rank v, ordered by m
// descending, then k.  Zero and normal inputs take the rounds; inf, NaN and
// subnormal inputs replay the rule level by level with running ratios
// (halving r_k is exact while it is normal, and below 1 it never splits).
// The host check tests/native/split_sequence_check.hip compares this with
// the level-by-level rule.
P2C_TUV_HD void tuv_split_sequence(const double inv_tol[3], uint64_t split[5], uint32_t* done_level) {
  uint64_t w0 = ~0ull, w1 = ~0ull, w2 = ~0ull, w3 = ~0ull, w4 = ~0ull;
  uint64_t cur = ~0ull;
  uint32_t level = 0;
  bool regular = true;
  int E[3], n[3];
  uint64_t mant[3];
  for (int k = 0; k < 3; ++k) {
    const uint64_t bits = __builtin_bit_cast(uint64_t, inv_tol[k]);
    const int bexp = (int)((bits >> 52) & 0x7FF);
    mant[k] = bits & ((1ull << 52) - 1ull);
    E[k] = bexp - 1023;
    const bool zero = (bits << 1) == 0;
    regular = regular && (zero || (bexp != 0 && bexp != 0x7FF)) && !(bits >> 63);
    int c = zero ? 0 : E[k] + (mant[k] != 0 ? 1 : 0);
    c = c < 0 ? 0 : (c > kTuvMaxExp ? kTuvMaxExp : c);
    n[k] = c;
  }
  if (regular) {
    // order: m descending, then k (a three-element sorting network)
    int o0 = 0, o1 = 1, o2 = 2;
    auto before = [&](int a, int b) { return mant[a] > mant[b] || (mant[a] == mant[b] && a < b); };
    if (before(o1, o0)) { const int t = o0; o0 = o1; o1 = t; }
    if (before(o2, o1)) { const int t = o1; o1 = o2; o2 = t; }
    if (before(o1, o0)) { const int t = o0; o0 = o1; o1 = t; }
    // dimension k splits in rounds [E_k - n_k + 1, E_k]; a round's codes
    // (the active dimensions in `ord`) are appended to the accumulator at
    // once, and a full 64-bit word is flushed to w0..w4
    const int k0 = o0, k1 = o1, k2 = o2;
    const int hi0 = get3(E, k0), lo0 = hi0 - get3(n, k0) + 1;
    const int hi1 = get3(E, k1), lo1 = hi1 - get3(n, k1) + 1;
    const int hi2 = get3(E, k2), lo2 = hi2 - get3(n, k2) + 1;
    int vmax = -100000, vmin = 100000;
    for (int k = 0; k < 3; ++k) {
      if (n[k] == 0) continue;
      vmax = E[k] > vmax ? E[k] : vmax;
      vmin = E[k] - n[k] + 1 < vmin ? E[k] - n[k] + 1 : vmin;
    }
    uint64_t acc = 0;  // codes of levels [64 * flushed / 2, level) not yet flushed
    uint32_t bits = 0, flushed = 0;
    for (int v = vmax; v >= vmin; --v) {
      uint64_t pat = 0;
      uint32_t pb = 0;
      if (v >= lo0 && v <= hi0) { pat |= (uint64_t)k0 << pb; pb += 2; }
      if (v >= lo1 && v <= hi1) { pat |= (uint64_t)k1 << pb; pb += 2; }
      if (v >= lo2 && v <= hi2) { pat |= (uint64_t)k2 << pb; pb += 2; }
      acc |= pat << bits;
      bits += pb;
      if (bits >= 64) {
        w0 = flushed == 0 ? acc : w0;
        w1 = flushed == 1 ? acc : w1;
        w2 = flushed == 2 ? acc : w2;
        w3 = flushed == 3 ? acc : w3;
        w4 = flushed == 4 ? acc : w4;
        ++flushed;
        bits -= 64;
        acc = bits > 0 ? pat >> (pb - bits) : 0ull;
      }
    }
    level = flushed * 32 + bits / 2;
    cur = acc | (bits >= 64 ? 0ull : (~0ull << bits));  // unused levels read 3 (no split)
  } else {
    double r0 = inv_tol[0], r1 = inv_tol[1], r2 = inv_tol[2];
    int e0 = 0, e1 = 0, e2 = 0;
    for (;;) {
      int d = 3;
      double br = 1.0;
      if (e0 < kTuvMaxExp && r0 > br) {
        br = r0;
        d = 0;
      }
      if (e1 < kTuvMaxExp && r1 > br) {
        br = r1;
        d = 1;
      }
      if (e2 < kTuvMaxExp && r2 > br) d = 2;
      if (d == 3) break;
      r0 = d == 0 ? r0 * 0.5 : r0;
      r1 = d == 1 ? r1 * 0.5 : r1;
      r2 = d == 2 ? r2 * 0.5 : r2;
      e0 += d == 0;
      e1 += d == 1;
      e2 += d == 2;
      split_append(cur, level, d, w0, w1, w2, w3, w4);
    }
  }
  if ((level & 31u) != 0) {
    const uint32_t wi = level >> 5;
    w0 = wi == 0 ? cur : w0;
    w1 = wi == 1 ? cur : w1;
    w2 = wi == 2 ? cur : w2;
    w3 = wi == 3 ? cur : w3;
    w4 = wi == 4 ? cur : w4;
  }
  split[0] = w0;
  split[1] = w1;
  split[2] = w2;
  split[3] = w3;
  split[4] = w4;
  *done_level = level;
}

// The positional constants of a pair from its 8 positions (rows 0-3 at
// t=0, 4-7 at t=1): start positions, displacements and the per-axis error
// bound.  Also what a compact (vertex-id) pair record is expanded with.
P2C_TUV_HD void tuv_setup_geometry(TuvPair* q, const double v[8][3], int is_ee, double ms) {
  q->is_ee = (uint32_t)is_ee;
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 3; ++k) {
      q->x0[i][k] = v[i][k];
      q->d[i][k] = v[4 + i][k] - v[i][k];
    }
  for (int k = 0; k < 3; ++k) {
    double m = 0.0;
    for (int i = 0; i < 8; ++i) {
      const double a = __builtin_fabs(v[i][k]);
      m = a > m ? a : m;
    }
    q->err[k] = kTuvErrScale * m + kTuvErrFloor + ms;
  }
}

// Sets the bits {start + p * r : 0 <= r < count} (p = 1, 2 or 3) of a
// 128-level mask; levels from 128 on are not kept here.
P2C_TUV_HD void set_periodic(uint64_t m[2], uint32_t start, uint32_t p, uint32_t count) {
  if (count == 0) return;
  const uint64_t base = p == 1 ? ~0ull : (p == 2 ? 0x5555555555555555ull : 0x9249249249249249ull);
  const uint32_t end = start + p * (count - 1u) + 1u;  // one past the last bit
#pragma unroll
  for (uint32_t w = 0; w < 2; ++w) {
    const uint32_t lo = 64u * w, hi = lo + 64u;
    if (end <= lo || start >= hi) continue;
    const uint32_t s0 = start > lo ? start - lo : 0u;
    const uint32_t e0 = end < hi ? end - lo : 64u;
    const uint32_t phase = start >= lo ? start - lo : (p - (lo - start) % p) % p;
    const uint64_t range = (e0 == 64u ? ~0ull : ((1ull << e0) - 1ull)) & ~((1ull << s0) - 1ull);
    m[w] |= (base << phase) & range;
  }
}

// The level masks, high codes and length of tuv_split_sequence's sequence,
// built without the 2-bit words: between two consecutive changes of the set
// of dimensions that split in a round (at most six), the rounds repeat one
// pattern of period p (the number of those dimensions), so each dimension's
// levels there are every p-th level from its place in the round.  Inputs the
// rounds do not cover (inf, NaN, subnormal) take tuv_split_sequence.  The
// host check tests/native/split_sequence_check.hip compares the two.
P2C_TUV_HD void tuv_split_masks_of(const double inv_tol[3], uint64_t tmask[2], uint64_t umask[2],
                                   uint64_t* split_hi, uint32_t* done_level) {
  bool regular = true;
  int E[3], n[3];
  uint64_t mant[3];
  for (int k = 0; k < 3; ++k) {
    const uint64_t bits = __builtin_bit_cast(uint64_t, inv_tol[k]);
    const int bexp = (int)((bits >> 52) & 0x7FF);
    mant[k] = bits & ((1ull << 52) - 1ull);
    E[k] = bexp - 1023;
    const bool zero = (bits << 1) == 0;
    regular = regular && (zero || (bexp != 0 && bexp != 0x7FF)) && !(bits >> 63);
    int c = zero ? 0 : E[k] + (mant[k] != 0 ? 1 : 0);
    c = c < 0 ? 0 : (c > kTuvMaxExp ? kTuvMaxExp : c);
    n[k] = c;
  }
  if (!regular) {
    uint64_t split[5];
    tuv_split_sequence(inv_tol, split, done_level);
    tuv_split_masks(split, tmask, umask, split_hi);
    return;
  }
  int o[3] = {0, 1, 2};  // m descending, then k
  auto before = [&](int a, int b) { return mant[a] > mant[b] || (mant[a] == mant[b] && a < b); };
  if (before(o[1], o[0])) { const int t = o[0]; o[0] = o[1]; o[1] = t; }
  if (before(o[2], o[1])) { const int t = o[1]; o[1] = o[2]; o[2] = t; }
  if (before(o[1], o[0])) { const int t = o[0]; o[0] = o[1]; o[1] = t; }
  int hi[3], lo[3];
  int vmax = -100000, vmin = 100000;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int k = o[i];
    const int nk = get3(n, k), ek = get3(E, k);
    hi[i] = nk > 0 ? ek : -100000;
    lo[i] = nk > 0 ? ek - nk + 1 : 100000;
    vmax = nk > 0 && ek > vmax ? ek : vmax;
    vmin = nk > 0 && lo[i] < vmin ? lo[i] : vmin;
  }
  tmask[0] = tmask[1] = umask[0] = umask[1] = 0;
  uint64_t codes_hi = ~0ull;  // levels 128..159, 3 = no split
  uint32_t level = 0;
  for (int v = vmax; v >= vmin;) {
    // the dimensions splitting in round v, in order, and the next round
    // below v at which that set changes
    int nxt = vmin - 1;
    uint32_t p = 0;
    int act[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (lo[i] <= v && v <= hi[i]) {
        act[p++] = o[i];
        nxt = lo[i] - 1 > nxt ? lo[i] - 1 : nxt;
      } else if (hi[i] < v && hi[i] > nxt) {
        nxt = hi[i];
      }
    }
    const uint32_t count = (uint32_t)(v - nxt);
    for (uint32_t i = 0; i < p; ++i) {
      const int k = act[i];
      if (k == 0) set_periodic(tmask, level + i, p, count);
      if (k == 1) set_periodic(umask, level + i, p, count);
    }
    if (level + p * count > kTuvMaskLevels) {  // past the masks: 2-bit codes (rare)
      for (uint32_t l = level > kTuvMaskLevels ? level : kTuvMaskLevels; l < level + p * count; ++l) {
        const uint32_t sh = (l - kTuvMaskLevels) * 2u;
        if (sh < 64u) codes_hi = (codes_hi & ~(3ull << sh)) | ((uint64_t)act[(l - level) % p] << sh);
      }
    }
    level += p * count;
    v = nxt;
  }
  *split_hi = codes_hi;
  *done_level = level;
}

// Per-pair constants from the 8 positions (rows 0-3 at t=0, 4-7 at t=1).
P2C_TUV_HD void tuv_setup(TuvPair* q, const double v[8][3], int is_ee, double ms, double tolerance) {
  tuv_setup_geometry(q, v, is_ee, ms);
  // unit-cube corner values (the operations of tuv_corner_range, one corner
  // at a time) and the largest change along each axis
  double F[2][2][2][3];
  for (int it = 0; it < 2; ++it)
    for (int k = 0; k < 3; ++k) {
      const double t = it;
      const double X0 = q->x0[0][k] + t * q->d[0][k];
      const double X1 = q->x0[1][k] + t * q->d[1][k];
      const double X2 = q->x0[2][k] + t * q->d[2][k];
      const double X3 = q->x0[3][k] + t * q->d[3][k];
      for (int iu = 0; iu < 2; ++iu)
        for (int iv = 0; iv < 2; ++iv) {
          const double u = iu, v = iv;
          if (!is_ee) {
            const double e1 = X2 - X1, e2 = X3 - X1;
            F[it][iu][iv][k] = X0 - ((X1 + u * e1) + v * e2);
          } else {
            F[it][iu][iv][k] = (X0 + u * (X1 - X0)) - (X2 + v * (X3 - X2));
          }
        }
    }
  double L[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int k = 0; k < 3; ++k) {
        const double dt = __builtin_fabs(F[1][a][b][k] - F[0][a][b][k]);
        const double du = __builtin_fabs(F[a][1][b][k] - F[a][0][b][k]);
        const double dv = __builtin_fabs(F[a][b][1][k] - F[a][b][0][k]);
        L[0] = dt > L[0] ? dt : L[0];
        L[1] = du > L[1] ? du : L[1];
        L[2] = dv > L[2] ? dv : L[2];
      }
  for (int k = 0; k < 3; ++k) q->inv_tol[k] = (3.0 * L[k]) / tolerance;
  tuv_split_masks_of(q->inv_tol, q->tmask, q->umask, &q->split_hi, &q->done_level);
}

// Box outcome: 0 = no zero inside, 1 = done (certifies a hit at its t_lo),
// 2 = undecided (split).  (tl, ul, vl) are the box's lower corner as exact
// doubles; its widths are 2^-e.
P2C_TUV_HD int tuv_classify_at(const TuvPair& q, double tl, double ul, double vl, const int e[3], uint32_t level,
                               double tolerance) {
  const double th = tl + __builtin_ldexp(1.0, -e[0]);
  const double uh = ul + __builtin_ldexp(1.0, -e[1]);
  const double vh = vl + __builtin_ldexp(1.0, -e[2]);
  double mn[3], mx[3];
  tuv_corner_range(q, tl, th, ul, uh, vl, vh, mn, mx);
  bool zero_in = true, box_in = true, small = true;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double eps = q.err[k];
    zero_in = zero_in && !(mn[k] > eps || mx[k] < -eps);
    box_in = box_in && !(mn[k] < -eps || mx[k] > eps);
    small = small && !(mx[k] - mn[k] > tolerance);
  }
  if (!zero_in) return 0;
  if (small || box_in || level >= q.done_level) return 1;
  return 2;
}

// The same from dyadic keys (lower corners scaled by 2^63).
P2C_TUV_HD int tuv_classify(const TuvPair& q, uint64_t kt, uint64_t ku, uint64_t kv, const int e[3],
                            uint32_t level, double tolerance) {
  return tuv_classify_at(q, key_to_f64(kt), key_to_f64(ku), key_to_f64(kv), e, level, tolerance);
}

// Exponents of the boxes at `level` (at most done_level): the splits of each
// dimension on levels 0..level-1, counted in the masks, then level by level
// past them.
P2C_TUV_HD void tuv_exponents(const TuvPair& q, uint32_t level, int e[3]) {
  const uint32_t lm = level < kTuvMaskLevels ? level : kTuvMaskLevels;
  const uint64_t m0 = lm >= 64 ? ~0ull : ((1ull << lm) - 1ull);
  const uint64_t m1 = lm >= 128 ? ~0ull : (lm > 64 ? ((1ull << (lm - 64)) - 1ull) : 0ull);
  int e0 = __builtin_popcountll(q.tmask[0] & m0) + __builtin_popcountll(q.tmask[1] & m1);
  int e1 = __builtin_popcountll(q.umask[0] & m0) + __builtin_popcountll(q.umask[1] & m1);
  int e2 = (int)lm - e0 - e1;
  for (uint32_t l = kTuvMaskLevels; l < level; ++l) {
    const int d = tuv_level_dim(q, l);
    e0 += d == 0;
    e1 += d == 1;
    e2 += d == 2;
  }
  e[0] = e0;
  e[1] = e1;
  e[2] = e2;
}

// Depth-first walk of the subtree rooted at box (kt,ku,kv) of level `root`,
// lower halves first, pruning boxes with t_lo >= *best.  Updates *best with
// the least t_lo of a done box, adds evaluated boxes to *boxes, and stops
// early (returning false) once *boxes exceeds `budget`.
P2C_TUV_HD bool tuv_walk(const TuvPair& q, uint64_t kt, uint64_t ku, uint64_t kv, uint32_t root,
                         double tolerance, uint64_t* best, uint64_t* boxes, uint64_t budget) {
  int e[3];
  tuv_exponents(q, root, e);
  uint64_t key[3] = {kt, ku, kv};
  double lo[3] = {key_to_f64(kt), key_to_f64(ku), key_to_f64(kv)};  // exact: <= 52 significant bits
  uint32_t level = root;
  bool visit = true;  // evaluate the current box
  for (;;) {
    if (visit) {
      int outcome = 0;
      const bool discard = (!q.is_ee && key[1] + key[2] > kOne63);
      if (!discard && key[0] < *best) {
        if (*boxes >= budget) return false;
        ++*boxes;
        outcome = tuv_classify_at(q, lo[0], lo[1], lo[2], e, level, tolerance);
        if (outcome == 1 && key[0] < *best) *best = key[0];
      }
      if (outcome == 2) {
        const int d = tuv_level_dim(q, level);
        put3(e, d, get3(e, d) + 1);
        ++level;  // lower child: corner unchanged
        continue;
      }
    }
    // next box in pre-order within the subtree: climb while an upper child
    visit = false;
    while (level > root) {
      const int d = tuv_level_dim(q, level - 1);
      const int ed = get3(e, d);
      const uint64_t bit = kOne63 >> ed;
      const double w = __builtin_ldexp(1.0, -ed);
      const uint64_t kd = get3(key, d);
      if (kd & bit) {
        put3(key, d, kd & ~bit);
        put3(lo, d, get3(lo, d) - w);
        put3(e, d, ed - 1);
        --level;
      } else {
        put3(key, d, kd | bit);  // upper sibling
        put3(lo, d, get3(lo, d) + w);
        visit = true;
        break;
      }
    }
    if (!visit) return true;
  }
}

// The stream kernel's walk: the same depth-first pre-order as tuv_walk (the
// same boxes, evaluated in the same order, pruned by the same rules), with
// the climb to the next box made in one jump instead of one level per move.
// low[] marks the levels 0..127 at which the path took the lower child while
// the upper sibling is still to be visited; an upper sibling outside the
// triangle (VF, u + v > 1) is never marked, since its whole subtree is
// skipped.  The next box after a finished subtree is the upper sibling at the
// deepest marked level; its exponents are population counts of the pair's
// level masks, and its keys keep the bits of the splits above it.
//
// Keys are 32-bit (lower corner * 2^31), so the walk takes the pairs whose
// dimensions split at most 31 times each and whose split sequence fits the
// masks (jump_walk_fits); a tolerance of 1e-6 needs extents near 1e3 to pass
// that.  Other pairs are handed to the wave pass unwalked.
constexpr uint32_t kOne31 = 1u << 31;
constexpr uint32_t kJumpNoHit = ~0u;

struct JumpWalk {
  uint32_t key[3];
  uint32_t best;  // least done t key, kJumpNoHit: none
  uint64_t low[2];
  uint32_t boxes;
  uint32_t ep;
  uint32_t level;
  bool ready;
};

P2C_TUV_HD bool jump_walk_fits(const TuvPair& q) {
  const uint32_t et = (uint32_t)(__builtin_popcountll(q.tmask[0]) + __builtin_popcountll(q.tmask[1]));
  const uint32_t eu = (uint32_t)(__builtin_popcountll(q.umask[0]) + __builtin_popcountll(q.umask[1]));
  return q.done_level <= kTuvMaskLevels && et <= 31u && eu <= 31u && q.done_level - et - eu <= 31u;
}

P2C_TUV_HD void start_jump_walk(JumpWalk& w) {
  w.key[0] = w.key[1] = w.key[2] = 0;
  w.best = kJumpNoHit;
  w.low[0] = w.low[1] = 0;
  w.boxes = 0;
  w.ep = 0;
  w.level = 0;
  w.ready = true;
}

// The 64-bit forms of a jump walk's keys and best (tuv_walk's encoding).
P2C_TUV_HD uint64_t jump_key64(uint32_t k) { return (uint64_t)k << 32; }
P2C_TUV_HD uint64_t jump_best64(uint32_t b) { return b == kJumpNoHit ? kTuvNoHit : jump_key64(b); }

// Moves to the lower child of a box that splits (the same lower corner, so it
// passes every skip test its parent passed) and marks the level when the
// upper sibling is to be visited later.
P2C_TUV_HD void jump_descend(const TuvPair& q, JumpWalk& w) {
  const uint32_t l = w.level;  // < done_level <= kTuvMaskLevels
  const int d = tuv_mask_dim(q, l);
  w.ep += 1u << (8 * d);
  const uint32_t half = kOne31 >> ((w.ep >> (8 * d)) & 0xffu);
  const uint32_t ku = w.key[1] + (d == 1 ? half : 0u), kv = w.key[2] + (d == 2 ? half : 0u);
  const bool visit = q.is_ee || ku + kv <= kOne31;  // keys below 2^31: no overflow
  const uint64_t m = visit ? (1ull << (l & 63u)) : 0ull;
  w.low[0] |= l < 64 ? m : 0ull;
  w.low[1] |= l < 64 ? 0ull : m;
  w.level = l + 1;
}

// jump_descend for a box below level 64 (its mark lands in low[0]).
P2C_TUV_HD void jump_descend_lo(const TuvPair& q, JumpWalk& w) {
  const uint32_t l = w.level;
  const uint32_t t = (uint32_t)(q.tmask[0] >> l) & 1u, u = (uint32_t)(q.umask[0] >> l) & 1u;
  const uint32_t d = t ? 0u : (u ? 1u : 2u);
  w.ep += 1u << (8 * d);
  const uint32_t half = kOne31 >> ((w.ep >> (8 * d)) & 0xffu);
  const uint32_t ku = w.key[1] + (d == 1 ? half : 0u), kv = w.key[2] + (d == 2 ? half : 0u);
  const bool visit = q.is_ee || ku + kv <= kOne31;
  w.low[0] |= visit ? (1ull << l) : 0ull;
  w.level = l + 1;
}

// The upper child at marked level l of the path through key[] (m0 / m1: the
// levels 0..l; st / su: level l splits t / u): its exponents by population
// count, its keys cut to the splits on levels 0..l (bit positions >= 31 - e)
// with the split at l taken upwards.
P2C_TUV_HD void jump_sibling(const TuvPair& q, const uint32_t key[3], uint32_t l, uint64_t m0, uint64_t m1, bool st,
                             bool su, uint32_t out[3], uint32_t* ep) {
  const uint32_t et = (uint32_t)(__builtin_popcountll(q.tmask[0] & m0) + __builtin_popcountll(q.tmask[1] & m1));
  const uint32_t eu = (uint32_t)(__builtin_popcountll(q.umask[0] & m0) + __builtin_popcountll(q.umask[1] & m1));
  const uint32_t ev = l + 1u - et - eu;
  const bool sv = !st && !su;
  const uint32_t s0 = 1u << (31u - et), s1 = 1u << (31u - eu), s2 = 1u << (31u - ev);
  out[0] = (key[0] | (st ? s0 : 0u)) & (0u - s0);
  out[1] = (key[1] | (su ? s1 : 0u)) & (0u - s1);
  out[2] = (key[2] | (sv ? s2 : 0u)) & (0u - s2);
  *ep = et | (eu << 8) | (ev << 16);
}

// Moves the walk to that upper child (ready unless pruned by best).
P2C_TUV_HD void jump_land(const TuvPair& q, JumpWalk& w, uint32_t l, uint64_t m0, uint64_t m1, bool st, bool su) {
  uint32_t k[3], ep;
  jump_sibling(q, w.key, l, m0, m1, st, su, k, &ep);
  w.key[0] = k[0];
  w.key[1] = k[1];
  w.key[2] = k[2];
  w.ep = ep;
  w.level = l + 1u;
  w.ready = w.key[0] < w.best;
}

// The upper child at the shallowest marked level, whose mark is cleared: the
// largest pending subtree of the walk, for another lane to take.  The walk
// must hold a mark.
P2C_TUV_HD void jump_take_shallowest(const TuvPair& q, JumpWalk& w, uint32_t out[3], uint32_t* ep, uint32_t* level) {
  const bool lo = w.low[0] != 0;
  const uint64_t x = lo ? w.low[0] : w.low[1];
  const uint32_t b = (uint32_t)__builtin_ctzll(x);
  const uint64_t bit = 1ull << b;
  w.low[0] = lo ? (x ^ bit) : w.low[0];
  w.low[1] = lo ? w.low[1] : (x ^ bit);
  const uint64_t upto = bit | (bit - 1ull);
  const bool st = ((lo ? q.tmask[0] : q.tmask[1]) & bit) != 0;
  const bool su = !st && ((lo ? q.umask[0] : q.umask[1]) & bit) != 0;
  const uint32_t l = b + (lo ? 0u : 64u);
  jump_sibling(q, w.key, l, lo ? upto : ~0ull, lo ? 0ull : upto, st, su, out, ep);
  *level = l + 1u;
}

// Whether level j < w.level of the walk's path took the lower child with an
// upper sibling to visit (not outside the triangle): the mark jump_descend
// would have left there.  The sibling (keys k, exponents *ep, level j + 1)
// is returned either way.  Rebuilds the marks of a walk resumed from its
// keys.
P2C_TUV_HD bool jump_sibling_at(const TuvPair& q, const JumpWalk& w, uint32_t j, uint32_t k[3], uint32_t* ep) {
  const bool hi = j >= 64;
  const uint64_t bit = 1ull << (j & 63u);
  const uint64_t upto = bit | (bit - 1ull);
  const bool st = ((hi ? q.tmask[1] : q.tmask[0]) & bit) != 0;
  const bool su = !st && ((hi ? q.umask[1] : q.umask[0]) & bit) != 0;
  jump_sibling(q, w.key, j, hi ? ~0ull : upto, hi ? upto : 0ull, st, su, k, ep);
  // the path's own key holds the bit of the split at j when it took the upper child
  const uint32_t d = st ? 0u : (su ? 1u : 2u);
  const uint32_t ed = (*ep >> (8 * d)) & 0xffu;
  const uint32_t own = d == 0 ? w.key[0] : (d == 1 ? w.key[1] : w.key[2]);
  const bool lower = (own & (1u << (31u - ed))) == 0;
  return lower && (q.is_ee || k[1] + k[2] <= kOne31);
}

P2C_TUV_HD bool jump_pending_at(const TuvPair& q, const JumpWalk& w, uint32_t j) {
  uint32_t k[3], ep;
  return jump_sibling_at(q, w, j, k, &ep);
}

// Moves to the next box after the current box's subtree; false when the walk
// is over.  The new box is ready unless its t_lo is no longer below best (then
// its subtree is skipped by the next jump).
P2C_TUV_HD bool jump_next(const TuvPair& q, JumpWalk& w) {
  const bool hi = w.low[1] != 0;
  const uint64_t x = hi ? w.low[1] : w.low[0];
  if (x == 0) return false;
  const uint32_t b = 63u - (uint32_t)__builtin_clzll(x);
  const uint64_t bit = 1ull << b;
  w.low[1] = hi ? (x ^ bit) : w.low[1];
  w.low[0] = hi ? w.low[0] : (x ^ bit);
  const uint64_t upto = bit | (bit - 1ull);
  const bool st = ((hi ? q.tmask[1] : q.tmask[0]) & bit) != 0;
  const bool su = !st && ((hi ? q.umask[1] : q.umask[0]) & bit) != 0;
  jump_land(q, w, b + (hi ? 64u : 0u), hi ? ~0ull : upto, hi ? upto : 0ull, st, su);
  return true;
}

// jump_next for a walk with no marks past level 63 (low[1] == 0).
P2C_TUV_HD bool jump_next_lo(const TuvPair& q, JumpWalk& w) {
  const uint64_t x = w.low[0];
  if (x == 0) return false;
  const uint32_t b = 63u - (uint32_t)__builtin_clzll(x);
  const uint64_t bit = 1ull << b;
  w.low[0] = x ^ bit;
  const bool st = (q.tmask[0] & bit) != 0;
  const bool su = !st && (q.umask[0] & bit) != 0;
  jump_land(q, w, b, bit | (bit - 1ull), 0ull, st, su);
  return true;
}

// Evaluates the box a jump walk stands on: 0 reject, 1 done, 2 split.  The
// keys' values are exact in FP64 (31 significant bits).
P2C_TUV_HD int jump_classify(const TuvPair& q, const JumpWalk& w, double tol) {
  const int e[3] = {(int)(w.ep & 0xffu), (int)((w.ep >> 8) & 0xffu), (int)((w.ep >> 16) & 0xffu)};
  return tuv_classify_at(q, (double)w.key[0] * 0x1p-31, (double)w.key[1] * 0x1p-31, (double)w.key[2] * 0x1p-31, e,
                         w.level, tol);
}

}  // namespace p2c
