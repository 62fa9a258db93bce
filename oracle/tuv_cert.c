/*
 * Oracle (test infrastructure only): (t,u,v) inclusion certificate in its
 * canonical sequential order.
 *
 * Method: the interval root finder of Wang, Ferguson, Schneider, Panozzo,
 * "A Large-Scale Benchmark and an Inclusion-Based Algorithm for Continuous
 * Collision Detection", ACM TOG 40(5), 2021 -- the library the reference
 * calls as ticcd::vertexFaceCCD / edgeEdgeCCD
 * (src/tools/tight_inclusion_full_query_benchmark.cpp:375-397, parameters
 * :24-40; src/tools/scene_object_envelope_native_ti_walltime.cpp:704-754).
 * That library is not vendored with the reference and not available here, so
 * bit parity with it is UNPINNED; this file states the method precisely and
 * is pinned by closed forms, exact-rational brute force and invariants
 * (tests/test_oracle_tuv.py).
 *
 * Definition implemented (shared, operation for operation, with the gfx950
 * kernels in p2c-ccd_amd/csrc/cert_tuv_math.h):
 *  F(t,u,v) = VF: p(t) - ((f0(t) + u e1(t)) + v e2(t)),  e_i = f_i - f0
 *             EE: (a0(t) + u ea(t)) - (b0(t) + v eb(t))
 *  with x(t) = x0 + t (x1 - x0).  F is multilinear, so over a box its range
 *  per coordinate lies between the min and max of its 8 corner values
 *  (IEEE minNum/maxNum: fmin / fmax, the same on the host and the device).
 *  err_k = 5.6e-15 * max|x_k| (+1e-300) bounds the FP64 error of one corner
 *  evaluation (first-order analysis gives 46 M eps for VF, 42 M eps for EE).
 *  A box is zero_in if every coordinate's corner range meets [-(err+ms), err+ms].
 *  A zero_in box is done if (a) every coordinate's corner range is <= tol,
 *  (b) it lies inside the eps box, or (c) no dimension is left to split.
 *  Boxes split along argmax_k width_k * inv_tol_k over dimensions with
 *  width_k * inv_tol_k > 1 and depth < 52 (ties: t, then u, then v);
 *  inv_tol_k = 3 L_k / tol with L_k the largest corner-to-corner change of F
 *  along axis k over the unit cube.  VF boxes with u_lo + v_lo > 1 are dropped.
 *  Result: hit iff some done box exists; toi = smallest t_lo over done boxes.
 *
 * Canonical order (defines the result when the work cap binds): best-first by
 * (t_lo, deeper first, u_lo, v_lo).  Before the first box with a new t_lo is
 * processed, if more than max_itr boxes with smaller t_lo were processed the
 * search stops: hit with toi = that t_lo (a conservative lower bound, since no
 * root lies below it).  A box count beyond 16 max_itr or a heap beyond
 * heap_cap stops the same way at the current t_lo.  Because the result
 * without a cap is the order-free "min t_lo over done boxes", and every box
 * with t_lo below that minimum is processed in any order, a parallel search
 * that processes at most max_itr boxes for a pair has exactly this result.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define TUV_CERR 5.6e-15
#define TUV_TINY 1e-300
#define TUV_MAXE 52
#define ONE63 (1ull << 63)

/* planted mistakes for tests/test_oracle_tuv.py: 1 drops the error bound,
 * 2 drops the VF u + v <= 1 restriction, 3 builds the second EE edge from the
 * wrong vertex */
int orc_tuv_mutation = 0;

typedef struct {
  double x0[4][3], d[4][3];
  int ee;
  double err[3];
  double ms;
  double tol;
  double inv_tol[3];
} Tuv;

/* corner evaluation: for t in {tl, th} (index it), u in {ul, uh}, v in {vl, vh} */
static void eval_box(const Tuv* q, double tl, double th, double ul, double uh, double vl, double vh,
                     double mn[3], double mx[3]) {
  const double ts[2] = {tl, th}, us[2] = {ul, uh}, vs[2] = {vl, vh};
  for (int k = 0; k < 3; ++k) {
    mn[k] = INFINITY;
    mx[k] = -INFINITY;
  }
  for (int it = 0; it < 2; ++it) {
    double X[4][3];
    for (int i = 0; i < 4; ++i)
      for (int k = 0; k < 3; ++k) X[i][k] = q->x0[i][k] + ts[it] * q->d[i][k];
    for (int iu = 0; iu < 2; ++iu)
      for (int iv = 0; iv < 2; ++iv)
        for (int k = 0; k < 3; ++k) {
          double f;
          if (!q->ee) {
            const double e1 = X[2][k] - X[1][k];
            const double e2 = X[3][k] - X[1][k];
            const double w1 = X[1][k] + us[iu] * e1;
            const double w = w1 + vs[iv] * e2;
            f = X[0][k] - w;
          } else {
            const double pa = X[0][k] + us[iu] * (X[1][k] - X[0][k]);
            const double pb = X[2][k] + vs[iv] * (X[3][k] - X[orc_tuv_mutation == 3 ? 1 : 2][k]);
            f = pa - pb;
          }
          mn[k] = fmin(mn[k], f);
          mx[k] = fmax(mx[k], f);
        }
  }
}

static void setup(Tuv* q, int is_ee, const double* v, double ms, double tol) {
  memset(q, 0, sizeof(*q));
  q->ee = is_ee;
  q->ms = ms;
  q->tol = tol;
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 3; ++k) {
      q->x0[i][k] = v[i * 3 + k];
      q->d[i][k] = v[(4 + i) * 3 + k] - v[i * 3 + k];
    }
  for (int k = 0; k < 3; ++k) {
    double m = 0.0;
    for (int i = 0; i < 8; ++i) {
      const double a = fabs(v[i * 3 + k]);
      if (a > m) m = a;
    }
    q->err[k] = orc_tuv_mutation == 1 ? 0.0 : TUV_CERR * m + TUV_TINY;
  }
  /* unit-cube corner values F[it][iu][iv][k] */
  double F[2][2][2][3];
  for (int it = 0; it < 2; ++it)
    for (int iu = 0; iu < 2; ++iu)
      for (int iv = 0; iv < 2; ++iv) {
        double mn[3], mx[3];
        eval_box(q, (double)it, (double)it, (double)iu, (double)iu, (double)iv, (double)iv, mn, mx);
        for (int k = 0; k < 3; ++k) F[it][iu][iv][k] = mn[k];
      }
  double L[3] = {0, 0, 0};
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int k = 0; k < 3; ++k) {
        const double dt = fabs(F[1][a][b][k] - F[0][a][b][k]);
        const double du = fabs(F[a][1][b][k] - F[a][0][b][k]);
        const double dv = fabs(F[a][b][1][k] - F[a][b][0][k]);
        if (dt > L[0]) L[0] = dt;
        if (du > L[1]) L[1] = du;
        if (dv > L[2]) L[2] = dv;
      }
  for (int k = 0; k < 3; ++k) q->inv_tol[k] = (3.0 * L[k]) / tol;
}

void orc_tuv_params(int is_ee, const double* v, double tolerance, double err_out[3], double inv_tol_out[3]) {
  Tuv q;
  setup(&q, is_ee, v, 0.0, tolerance);
  for (int k = 0; k < 3; ++k) {
    err_out[k] = q.err[k];
    inv_tol_out[k] = q.inv_tol[k];
  }
}

/* corner range of F over one box, and the pair's error bound (for the
 * enclosure test) */
void orc_tuv_box(int is_ee, const double* v, double tl, double th, double ul, double uh, double vl, double vh,
                 double mn[3], double mx[3], double err[3]) {
  Tuv q;
  setup(&q, is_ee, v, 0.0, 1e-6);
  eval_box(&q, tl, th, ul, uh, vl, vh, mn, mx);
  for (int k = 0; k < 3; ++k) err[k] = q.err[k];
}

/* split dimension for exponents e[3], or -1 when no dimension may split */
static int split_dim(const Tuv* q, const int e[3]) {
  int best = -1;
  double br = 1.0;
  for (int k = 0; k < 3; ++k) {
    if (e[k] >= TUV_MAXE) continue;
    const double r = ldexp(1.0, -e[k]) * q->inv_tol[k];
    if (r > br) {
      br = r;
      best = k;
    }
  }
  return best;
}

typedef struct {
  uint64_t kt, ku, kv; /* lower corners scaled by 2^63 */
  uint8_t level, e[3];
} Box;

/* strict total order: t_lo, deeper first, u_lo, v_lo */
static int box_less(const Box* a, const Box* b) {
  if (a->kt != b->kt) return a->kt < b->kt;
  if (a->level != b->level) return a->level > b->level;
  if (a->ku != b->ku) return a->ku < b->ku;
  return a->kv < b->kv;
}

static void heap_push(Box* h, int64_t* n, Box b) {
  int64_t i = (*n)++;
  while (i > 0) {
    const int64_t p = (i - 1) / 2;
    if (!box_less(&b, &h[p])) break;
    h[i] = h[p];
    i = p;
  }
  h[i] = b;
}

static Box heap_pop(Box* h, int64_t* n) {
  Box top = h[0];
  const Box last = h[--(*n)];
  int64_t i = 0;
  for (;;) {
    int64_t c = 2 * i + 1;
    if (c >= *n) break;
    if (c + 1 < *n && box_less(&h[c + 1], &h[c])) ++c;
    if (!box_less(&h[c], &last)) break;
    h[i] = h[c];
    i = c;
  }
  if (*n > 0) h[i] = last;
  return top;
}

static double key_to_double(uint64_t k) { return ldexp((double)k, -63); }

int orc_tuv_eval(int is_ee, const double* v, double ms, double tolerance, double t_max, int64_t max_itr,
                 int64_t heap_cap, orc_tuv_result* out) {
  if (t_max != 1.0 || !(tolerance > 0.0) || !(ms >= 0.0) || max_itr <= 0 || heap_cap < 8) return -1;
  for (int i = 0; i < 24; ++i)
    if (!isfinite(v[i])) return -4;
  Tuv q;
  setup(&q, is_ee, v, ms, tolerance);
  int64_t hcap = heap_cap < 256 ? heap_cap : 256; /* grows by doubling up to heap_cap */
  Box* heap = (Box*)malloc(sizeof(Box) * (size_t)hcap);
  if (!heap) return -5;
  int64_t hn = 0;
  Box root = {0, 0, 0, 0, {0, 0, 0}};
  heap_push(heap, &hn, root);
  uint64_t best = UINT64_MAX;
  uint64_t last_t = UINT64_MAX;
  int64_t popped = 0;
  int capped = 0;
  uint64_t cap_key = 0;
  const int64_t hard_cap = 16 * max_itr;
  while (hn > 0) {
    const Box b = heap_pop(heap, &hn);
    if (b.kt >= best) break;
    if (b.kt != last_t) {
      last_t = b.kt;
      if (popped > max_itr) {
        capped = 1;
        cap_key = b.kt;
        break;
      }
    }
    if (popped >= hard_cap) {
      capped = 1;
      cap_key = b.kt;
      break;
    }
    ++popped;
    const int e[3] = {b.e[0], b.e[1], b.e[2]};
    const double tl = key_to_double(b.kt), th = key_to_double(b.kt + (ONE63 >> e[0]));
    const double ul = key_to_double(b.ku), uh = key_to_double(b.ku + (ONE63 >> e[1]));
    const double vl = key_to_double(b.kv), vh = key_to_double(b.kv + (ONE63 >> e[2]));
    double mn[3], mx[3];
    eval_box(&q, tl, th, ul, uh, vl, vh, mn, mx);
    int zero_in = 1, box_in = 1, small = 1;
    for (int k = 0; k < 3; ++k) {
      const double eps = q.err[k] + q.ms;
      if (mn[k] > eps || mx[k] < -eps) zero_in = 0;
      if (mn[k] < -eps || mx[k] > eps) box_in = 0;
      if (mx[k] - mn[k] > q.tol) small = 0;
    }
    if (!zero_in) continue;
    const int d = split_dim(&q, e);
    if (small || box_in || d < 0) {
      if (b.kt < best) best = b.kt;
      continue;
    }
    Box lo = b, hi = b;
    lo.level = hi.level = (uint8_t)(b.level + 1);
    lo.e[d] = hi.e[d] = (uint8_t)(e[d] + 1);
    const uint64_t half = ONE63 >> (e[d] + 1);
    if (d == 0) hi.kt += half;
    else if (d == 1) hi.ku += half;
    else hi.kv += half;
    const Box ch[2] = {lo, hi};
    for (int c = 0; c < 2; ++c) {
      if (!is_ee && orc_tuv_mutation != 2 && ch[c].ku + ch[c].kv > ONE63) continue; /* u_lo + v_lo > 1 */
      if (hn >= hcap) {
        if (hcap >= heap_cap) {
          capped = 1;
          cap_key = b.kt;
          goto done;
        }
        hcap = hcap * 2 < heap_cap ? hcap * 2 : heap_cap;
        Box* grown = (Box*)realloc(heap, sizeof(Box) * (size_t)hcap);
        if (!grown) {
          free(heap);
          return -5;
        }
        heap = grown;
      }
      heap_push(heap, &hn, ch[c]);
    }
  }
done:
  free(heap);
  out->boxes = popped;
  out->capped = capped;
  if (capped) {
    out->hit = 1;
    out->toi = key_to_double(cap_key);
  } else if (best != UINT64_MAX) {
    out->hit = 1;
    out->toi = key_to_double(best);
  } else {
    out->hit = 0;
    out->toi = INFINITY;
  }
  return 0;
}

int orc_tuv_eval_batch(int is_ee, const double* v, int64_t n, double ms, double tolerance, double t_max,
                       int64_t max_itr, int64_t heap_cap, orc_tuv_result* out, int nthreads) {
  int rc = 0;
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads) reduction(min : rc)
  for (int64_t i = 0; i < n; ++i) {
    const int r = orc_tuv_eval(is_ee, v + i * 24, ms, tolerance, t_max, max_itr, heap_cap, &out[i]);
    if (r < rc) rc = r;
  }
  return rc;
}
