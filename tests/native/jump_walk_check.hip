// Host check: the stream kernel's jump walk (one jump per climb, pending
// upper siblings marked per level, exponents from the level masks) against
// tuv_walk, the plain depth-first walk it restates: the same number of boxes
// evaluated and the same least done t_lo, on random vertex-face and
// edge-edge pairs -- uniform, near contact (hits and deep walks), coplanar,
// and scaled so that walks run past 64 levels.  Pairs the jump walk does not
// take (jump_walk_fits) are left to the wave pass by the kernel and are
// skipped here; scale 1e4 makes some.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include "cert_tuv_math.h"
using namespace p2c;

static bool jump_walk(const TuvPair& q, double tol, uint64_t* best, uint64_t* boxes, uint64_t budget) {
  JumpWalk w;
  start_jump_walk(w);
  for (;;) {
    if (w.ready) {
      if (w.boxes >= budget) {
        *best = jump_best64(w.best);
        *boxes = w.boxes;
        return false;
      }
      ++w.boxes;
      const int outcome = jump_classify(q, w, tol);
      if (outcome == 2) {
        if (w.level >= 64) jump_descend(q, w);
        else jump_descend_lo(q, w);
        continue;
      }
      if (outcome == 1 && w.key[0] < w.best) w.best = w.key[0];
      w.ready = false;
    }
    if (!(w.low[1] != 0 ? jump_next(q, w) : jump_next_lo(q, w))) break;
  }
  *best = jump_best64(w.best);
  *boxes = w.boxes;
  return true;
}

// The wave pass's use of the same walk on the host: a walk that now and then
// gives its shallowest pending subtree to a stack (as lanes do for idle
// lanes), each stacked subtree walked by a fresh walk rooted at it; the
// least done t_lo must equal the plain walk's.
static bool donating_walk(const TuvPair& q, double tol, std::mt19937_64& rng, uint64_t* best) {
  struct Box { uint32_t k[3], ep, level; };
  Box stack[4096];
  int sp = 0;
  uint32_t bst = kJumpNoHit;
  JumpWalk w;
  start_jump_walk(w);
  bool active = true;
  uint64_t steps = 0;
  while (active || sp > 0) {
    if (!active) {
      const Box b = stack[--sp];
      start_jump_walk(w);
      w.key[0] = b.k[0]; w.key[1] = b.k[1]; w.key[2] = b.k[2];
      w.ep = b.ep; w.level = b.level; w.best = bst;
      w.ready = w.key[0] < w.best;
      active = true;
    }
    if (++steps > 4000000) return false;
    bool finished = !w.ready;
    if (w.ready) {
      const int outcome = jump_classify(q, w, tol);
      if (outcome == 2) jump_descend(q, w);
      else {
        finished = true;
        if (outcome == 1 && w.key[0] < w.best) w.best = w.key[0];
      }
    }
    bst = w.best < bst ? w.best : bst;
    w.best = bst;
    if (finished && !jump_next(q, w)) active = false;
    if (active && (rng() % 7) == 0 && (w.low[0] | w.low[1]) != 0 && sp < 4096) {
      Box b;
      jump_take_shallowest(q, w, b.k, &b.ep, &b.level);
      stack[sp++] = b;
    }
  }
  *best = jump_best64(bst);
  return true;
}

// Marks rebuilt from the keys (jump_pending_at) against the marks a walk from
// the root carries, at every step of the walk.
static long rebuild_check(const TuvPair& q, double tol) {
  JumpWalk w;
  start_jump_walk(w);
  long bad = 0, steps = 0;
  for (;;) {
    if (++steps > 200000) break;
    uint64_t m[2] = {0, 0};
    for (uint32_t j = 0; j < w.level; ++j)
      if (jump_pending_at(q, w, j)) m[j >> 6] |= 1ull << (j & 63u);
    if (m[0] != w.low[0] || m[1] != w.low[1]) ++bad;
    if (w.ready) {
      const int outcome = jump_classify(q, w, tol);
      if (outcome == 2) {
        jump_descend(q, w);
        continue;
      }
      if (outcome == 1 && w.key[0] < w.best) w.best = w.key[0];
      w.ready = false;
    }
    if (!jump_next(q, w)) break;
  }
  return bad;
}

int main() {
  std::mt19937_64 rng(11);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  std::normal_distribution<double> N(0.0, 1.0);
  const double tol = 1e-6;
  const uint64_t budget = 20000;
  long n = 0, bad = 0, hits = 0, deep = 0, skipped = 0, over = 0, donated = 0, rebuilt = 0;
  uint64_t total_boxes = 0;
  for (int i = 0; i < 120000; ++i) {
    const int is_ee = i & 1;
    const int mode = (i >> 1) % 5;
    double v[8][3];
    for (int r = 0; r < 4; ++r)
      for (int k = 0; k < 3; ++k) {
        v[r][k] = U(rng);
        v[4 + r][k] = v[r][k] + 0.1 * N(rng);
      }
    if (mode == 1 || mode == 2) {
      // near contact: shift the first vertex (VF) or edge (EE) so that its
      // point at a random (u, v) meets the other primitive at a random t,
      // up to a tiny gap
      const double t = 0.5 * (U(rng) + 1.0), a = 0.5 * (U(rng) + 1.0), b = 0.5 * (U(rng) + 1.0);
      const double gap = mode == 1 ? 1e-8 * N(rng) : 0.0;
      double p[3], s[3];
      for (int k = 0; k < 3; ++k) {
        double x[4];
        for (int r = 0; r < 4; ++r) x[r] = v[r][k] + t * (v[4 + r][k] - v[r][k]);
        if (!is_ee) {
          const double uu = a * (1.0 - b), vv = b * a;  // inside the triangle
          p[k] = x[1] + uu * (x[2] - x[1]) + vv * (x[3] - x[1]);
          s[k] = x[0];
        } else {
          p[k] = x[2] + b * (x[3] - x[2]);
          s[k] = x[0] + a * (x[1] - x[0]);
        }
      }
      for (int k = 0; k < 3; ++k) {
        const double sh = p[k] - s[k] + (k == 2 ? gap : 0.0);
        v[0][k] += sh;
        v[4][k] += sh;
        if (is_ee) {
          v[1][k] += sh;
          v[5][k] += sh;
        }
      }
    } else if (mode == 3) {
      // coplanar: every vertex in z = 0 at both ends
      for (int r = 0; r < 8; ++r) v[r][2] = 0.0;
    } else if (mode == 4) {
      const double scale = (i / 10) % 7 == 0 ? 1e4 : ((i % 3 == 0) ? 40.0 : 4.0);  // deeper split sequences
      for (int r = 0; r < 8; ++r)
        for (int k = 0; k < 3; ++k) v[r][k] *= scale;
    }
    TuvPair q;
    tuv_setup(&q, v, is_ee, 0.0, tol);
    if (!jump_walk_fits(q)) {
      ++skipped;
      continue;
    }
    uint64_t b0 = kTuvNoHit, n0 = 0, b1 = kTuvNoHit, n1 = 0;
    const bool c0 = tuv_walk(q, 0, 0, 0, 0, tol, &b0, &n0, budget);
    const bool c1 = jump_walk(q, tol, &b1, &n1, budget);
    ++n;
    if (c0 && (i % 4) == 0) {
      uint64_t b2 = kTuvNoHit;
      if (!donating_walk(q, tol, rng, &b2) || b2 != b0) {
        if (bad < 5) printf("donating walk mismatch pair %d: %llu vs %llu\n", i, (unsigned long long)b2,
                            (unsigned long long)b0);
        ++bad;
      }
      ++donated;
    }
    if (c0 && (i % 16) == 0 && n0 < 20000) {
      const long rb = rebuild_check(q, tol);
      if (rb) {
        if (bad < 5) printf("mark rebuild mismatch pair %d: %ld steps\n", i, rb);
        ++bad;
      }
      ++rebuilt;
    }
    if (!c0) ++over;
    if (q.done_level > 64) ++deep;
    if (b0 != kTuvNoHit) ++hits;
    total_boxes += n0;
    if (c0 != c1 || b0 != b1 || n0 != n1) {
      if (bad < 5)
        printf("mismatch pair %d (ee %d, mode %d): plain best %llu boxes %llu, jump best %llu boxes %llu\n", i, is_ee,
               mode, (unsigned long long)b0, (unsigned long long)n0, (unsigned long long)b1,
               (unsigned long long)n1);
      ++bad;
    }
  }
  printf("jump_walk_check: %ld pairs (%ld hits, %ld past level 64, %ld over budget, %llu boxes; %ld not taken by the "
         "jump walk, skipped; %ld walked with donations, %ld with rebuilt marks), %ld mismatches\n",
         n, hits, deep, over, (unsigned long long)total_boxes, skipped, donated, rebuilt, bad);
  return bad != 0 || hits == 0 || deep == 0 || skipped == 0;
}
