/*
 * Oracle (test infrastructure only): CPU restatement of the reference's
 * interval certificate, src/cpp/certificate/certificate_engine.cpp, and of
 * its CUDA batch twin's walk, src/cpp/certificate/certificate_cuda.cu:286-346
 * (orc_ref_eval_twin).
 *
 * Written as a plain recursive C program, mirroring the reference's control
 * flow (the shipped HIP kernel uses a different, stackless traversal), so that
 * agreement between the two is evidence and not a copy checking itself.
 * Compiled with -ffp-contract=off; every expression keeps the reference's
 * operand order so FP64 results are bit-identical.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

static const double kEps = 1.0e-14; /* certificate_engine.cpp:20 */

/* Planted mistakes for the tests that show the golden vectors have teeth
 * (0 = faithful): 1 right half first, 2 motion radius drops the point term,
 * 3 lerp written p0 + t (p1 - p0), 4 margin keeps eps_space. */
int orc_ref_mutation = 0;

/* std::max / std::min / std::clamp semantics (NaN behaviour included). */
static double smax(double a, double b) { return (a < b) ? b : a; }
static double smin(double a, double b) { return (b < a) ? b : a; }
static double sclamp(double v, double lo, double hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }

typedef struct { double x, y, z; } V3;

static V3 v_add(V3 a, V3 b) { V3 r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static V3 v_sub(V3 a, V3 b) { V3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static V3 v_scale(V3 a, double s) { V3 r = {a.x * s, a.y * s, a.z * s}; return r; }
static double v_dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static V3 v_cross(V3 a, V3 b) {
  V3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
  return r;
}
static double v_norm(V3 a) { return sqrt(v_dot(a, a)); }
static V3 v_load(const double* p) { V3 r = {p[0], p[1], p[2]}; return r; }

/* PositionAt, certificate_engine.cpp:67-69 */
static V3 pos_at(const orc_traj* tr, double t) {
  if (orc_ref_mutation == 3)
    return v_add(v_load(tr->p0), v_scale(v_sub(v_load(tr->p1), v_load(tr->p0)), t));
  return v_add(v_scale(v_load(tr->p0), 1.0 - t), v_scale(v_load(tr->p1), t));
}

/* MaxDisplacementFromMidpoint, :71-76 */
static double max_disp(const orc_traj* tr, double t0, double t1) {
  const double len = smax(0.0, t1 - t0);
  return 0.5 * len * v_norm(v_sub(v_load(tr->p1), v_load(tr->p0)));
}

static double dist_pp(V3 a, V3 b) { return v_norm(v_sub(a, b)); } /* :78-80 */

/* DistancePointSegment, :82-90 */
static double dist_ps(V3 p, V3 a, V3 b) {
  const V3 ab = v_sub(b, a);
  const double ab2 = v_dot(ab, ab);
  if (ab2 <= kEps) return dist_pp(p, a);
  const double t = sclamp(v_dot(v_sub(p, a), ab) / ab2, 0.0, 1.0);
  return dist_pp(p, v_add(a, v_scale(ab, t)));
}

/* Denominator guard of the segment-segment distance: 0 = the C++ engine's
 * `denom != 0.0`, 1 = the Python twin's `abs(denom) > 1e-14`
 * (pure_exact_cpu.py:126); set only around a Python-twin walk. */
static int g_py_guard = 0;

/* DistanceSegmentSegment, :92-131 */
static double dist_ss(V3 p1, V3 q1, V3 p2, V3 q2) {
  const V3 d1 = v_sub(q1, p1);
  const V3 d2 = v_sub(q2, p2);
  const V3 r = v_sub(p1, p2);
  const double a = v_dot(d1, d1);
  const double e = v_dot(d2, d2);
  const double f = v_dot(d2, r);
  double s = 0.0, t = 0.0;
  if (a <= kEps && e <= kEps) return dist_pp(p1, p2);
  if (a <= kEps) {
    t = sclamp(f / e, 0.0, 1.0);
  } else {
    const double c = v_dot(d1, r);
    if (e <= kEps) {
      s = sclamp(-c / a, 0.0, 1.0);
    } else {
      const double b = v_dot(d1, d2);
      const double denom = a * e - b * b;
      if (g_py_guard ? fabs(denom) > 1.0e-14 : denom != 0.0) s = sclamp((b * f - c * e) / denom, 0.0, 1.0);
      t = (b * s + f) / e;
      if (t < 0.0) {
        t = 0.0;
        s = sclamp(-c / a, 0.0, 1.0);
      } else if (t > 1.0) {
        t = 1.0;
        s = sclamp((b - c) / a, 0.0, 1.0);
      }
    }
  }
  return dist_pp(v_add(p1, v_scale(d1, s)), v_add(p2, v_scale(d2, t)));
}

/* DistancePointTriangle, :133-184 (Voronoi regions; degenerate fallback) */
static double dist_pt(V3 p, V3 a, V3 b, V3 c) {
  const V3 ab = v_sub(b, a);
  const V3 ac = v_sub(c, a);
  const V3 n = v_cross(ab, ac);
  if (v_dot(n, n) <= kEps) {
    /* std::min({x, y, z}) keeps the first smallest */
    double m = dist_ps(p, a, b);
    const double y = dist_ps(p, b, c);
    const double z = dist_ps(p, c, a);
    if (y < m) m = y;
    if (z < m) m = z;
    return m;
  }
  const V3 ap = v_sub(p, a);
  const double d1 = v_dot(ab, ap);
  const double d2 = v_dot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return dist_pp(p, a);
  const V3 bp = v_sub(p, b);
  const double d3 = v_dot(ab, bp);
  const double d4 = v_dot(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) return dist_pp(p, b);
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
    const double v = d1 / (d1 - d3);
    return dist_pp(p, v_add(a, v_scale(ab, v)));
  }
  const V3 cp = v_sub(p, c);
  const double d5 = v_dot(ab, cp);
  const double d6 = v_dot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return dist_pp(p, c);
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
    const double w = d2 / (d2 - d6);
    return dist_pp(p, v_add(a, v_scale(ac, w)));
  }
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return dist_pp(p, v_add(b, v_scale(v_sub(c, b), w)));
  }
  return fabs(v_dot(ap, n) / v_norm(n));
}

double orc_pt_distance(const orc_prim* pr, double t) { /* :186-191 */
  return dist_pt(pos_at(&pr->v[0], t), pos_at(&pr->v[1], t), pos_at(&pr->v[2], t),
                 pos_at(&pr->v[3], t));
}

double orc_ee_distance(const orc_prim* pr, double t) { /* :193-198 */
  return dist_ss(pos_at(&pr->v[0], t), pos_at(&pr->v[1], t), pos_at(&pr->v[2], t),
                 pos_at(&pr->v[3], t));
}

/* PointTriangleMotionRadius / EdgeEdgeMotionRadius, :200-222 */
static double motion_radius(int family, const orc_prim* pr, double t0, double t1) {
  if (family == 1) {
    const double pd = max_disp(&pr->v[0], t0, t1);
    double td = max_disp(&pr->v[1], t0, t1); /* std::max({..}) keeps first largest */
    const double y = max_disp(&pr->v[2], t0, t1);
    const double z = max_disp(&pr->v[3], t0, t1);
    if (td < y) td = y;
    if (td < z) td = z;
    if (orc_ref_mutation == 2) return td;
    return pd + td;
  }
  const double ea = smax(max_disp(&pr->v[0], t0, t1), max_disp(&pr->v[1], t0, t1));
  const double eb = smax(max_disp(&pr->v[2], t0, t1), max_disp(&pr->v[3], t0, t1));
  return ea + eb;
}

typedef struct {
  int family;
  const orc_prim* pr;
  double eps_time, eps_space;
  int max_depth;
  int nodes;
} Ctx;

static double dist(const Ctx* c, double t) {
  return c->family == 1 ? orc_pt_distance(c->pr, t) : orc_ee_distance(c->pr, t);
}

/* MakePrimitiveResult, :266-277 */
static orc_result make_result(const Ctx* c, int status, double t0, double t1) {
  orc_result r;
  memset(&r, 0, sizeof(r));
  r.status = status;
  r.covered_feature_mask = status == 1 ? (uint32_t)c->family : 0u;
  r.interval_t0 = t0;
  r.interval_t1 = t1;
  r.toi_upper = t1;
  r.safe_margin_lb = 0.0;
  r.witness_family = 0;
  r.witness_id_a = -1;
  r.witness_id_b = -1;
  r.reason_code = 0;
  r.next_refinement_mode = 0;
  return r;
}

/* EvaluateIntervalRecursive, :300-386 */
static orc_result eval_rec(Ctx* c, double t0, double t1, int depth) {
  c->nodes++;
  const double mid = 0.5 * (t0 + t1);
  const double st[3] = {t0, mid, t1};
  const double sd[3] = {dist(c, t0), dist(c, mid), dist(c, t1)};
  for (int i = 0; i < 3; ++i) {
    if (sd[i] <= c->eps_space) {
      orc_result r = make_result(c, 0, t0, t1);
      r.toi_upper = st[i];
      r.witness_family = c->family;
      r.witness_id_a = c->pr->id_a;
      r.witness_id_b = c->pr->id_b;
      r.safe_margin_lb = 0.0;
      return r;
    }
  }
  /* lower_bound_fn: distance_fn(mid) - motion radius (:500-502, :536-538) */
  const double lb = dist(c, mid) - motion_radius(c->family, c->pr, t0, t1);
  if (lb > c->eps_space) {
    orc_result r = make_result(c, 1, t0, t1);
    r.safe_margin_lb = smax(0.0, orc_ref_mutation == 4 ? lb : lb - c->eps_space);
    return r;
  }
  if (depth >= c->max_depth || t1 - t0 <= c->eps_time) {
    orc_result r = make_result(c, 2, t0, t1);
    r.reason_code = 2;
    r.next_refinement_mode = 1;
    return r;
  }
  if (orc_ref_mutation == 1) {
    orc_result right = eval_rec(c, mid, t1, depth + 1);
    if (right.status == 0) return right;
    orc_result left = eval_rec(c, t0, mid, depth + 1);
    if (left.status == 0) return left;
    orc_result r = make_result(c, (left.status == 1 && right.status == 1) ? 1 : 2, t0, t1);
    if (r.status == 1) r.safe_margin_lb = smin(left.safe_margin_lb, right.safe_margin_lb);
    else { r.reason_code = 2; r.next_refinement_mode = 1; }
    return r;
  }
  orc_result left = eval_rec(c, t0, mid, depth + 1);
  if (left.status == 0) return left;
  orc_result right = eval_rec(c, mid, t1, depth + 1);
  if (right.status == 0) return right;
  if (left.status == 1 && right.status == 1) {
    orc_result r = make_result(c, 1, t0, t1);
    r.safe_margin_lb = smin(left.safe_margin_lb, right.safe_margin_lb);
    return r;
  }
  orc_result r = make_result(c, 2, t0, t1);
  r.reason_code = 2;
  r.next_refinement_mode = 1;
  return r;
}

static int finite_traj(const orc_traj* t) {
  for (int k = 0; k < 3; ++k)
    if (!isfinite(t->p0[k]) || !isfinite(t->p1[k])) return 0;
  return 1;
}

/* EvaluatePointTriangleInterval / EvaluateEdgeEdgeInterval, :481-551, with
 * ValidateConfig :224-232, ValidateInterval :234-242, primitive checks :244-264 */
int orc_ref_eval(int family, const orc_prim* prim, double t0, double t1, double eps_time,
                 double eps_space, int max_depth, orc_result* out) {
  if (!isfinite(eps_time) || eps_time <= 0.0) return -1;
  if (!isfinite(eps_space) || eps_space <= 0.0) return -1;
  if (!isfinite(t0) || !isfinite(t1)) return -2;
  if (t0 < 0.0 || t1 > 1.0 || t0 > t1) return -2;
  if (prim->id_a < 0 || prim->id_b < 0) return -3;
  for (int i = 0; i < 4; ++i)
    if (!finite_traj(&prim->v[i])) return -4;
  Ctx c = {family, prim, eps_time, eps_space, max_depth, 0};
  *out = eval_rec(&c, t0, t1, 0);
  out->nodes = c.nodes;
  return 0;
}

/* EvaluateIntervalIterative of the CUDA twin (certificate_cuda.cu:286-346):
 * an explicit stack of at most 128 intervals, right half pushed under the
 * left.  The first popped interval that touches (collision) or that may not
 * be split (depth limit, eps_time, or a stack that two more entries would
 * fill) ends the walk; where the CPU engine records an undecided leaf and
 * keeps searching, this walk returns UNDECIDED on that leaf's interval.
 * Separated intervals contribute max(0, lb - eps_space) to a running
 * minimum that starts at 1e300.  `nodes` counts popped intervals.  The
 * distance and motion-radius functions are the CPU engine's (the twin's are
 * the same formulas; nvcc's default FMA contraction may change their last
 * bits there). */
static orc_result eval_twin(Ctx* c, double T0, double T1) {
  double s0[128], s1[128];
  int sd[128];
  int size = 0;
  s0[size] = T0; s1[size] = T1; sd[size] = 0; ++size;
  double safe = 1.0e300;
  while (size > 0) {
    --size;
    const double t0 = s0[size], t1 = s1[size];
    const int depth = sd[size];
    c->nodes++;
    const double mid = 0.5 * (t0 + t1);
    const double d0 = dist(c, t0), dm = dist(c, mid), d1 = dist(c, t1);
    if (d0 <= c->eps_space || dm <= c->eps_space || d1 <= c->eps_space) {
      orc_result r = make_result(c, 0, t0, t1);
      r.toi_upper = d0 <= c->eps_space ? t0 : (dm <= c->eps_space ? mid : t1);
      r.witness_family = c->family;
      r.witness_id_a = c->pr->id_a;
      r.witness_id_b = c->pr->id_b;
      return r;
    }
    const double lb = dm - motion_radius(c->family, c->pr, t0, t1);
    if (lb > c->eps_space) {
      safe = fmin(safe, fmax(0.0, lb - c->eps_space));
      continue;
    }
    if (depth >= c->max_depth || t1 - t0 <= c->eps_time || size + 2 >= 128) {
      orc_result r = make_result(c, 2, t0, t1);
      r.reason_code = 2;
      r.next_refinement_mode = 1;
      return r;
    }
    s0[size] = mid; s1[size] = t1; sd[size] = depth + 1; ++size;
    s0[size] = t0; s1[size] = mid; sd[size] = depth + 1; ++size;
  }
  orc_result r = make_result(c, 1, T0, T1);
  r.safe_margin_lb = safe < 1.0e299 ? safe : 0.0;
  return r;
}

int orc_ref_eval_twin_batch(int family, const orc_prim* prims, int64_t n, double t0, double t1,
                            double eps_time, double eps_space, int max_depth, orc_result* out) {
  for (int64_t i = 0; i < n; ++i) {
    /* The reference's RunCudaBatch (certificate_cuda.cu:456-552) validates
       neither config, interval nor primitives; the library refuses them as
       the CPU engine does, a deliberate difference, and this restatement
       validates the same way so both agree on what is refused. */
    orc_result probe;
    const int rc = orc_ref_eval(family, &prims[i], t0, t1, eps_time, eps_space, 0, &probe);
    if (rc != 0) return rc;
    Ctx c = {family, &prims[i], eps_time, eps_space, max_depth, 0};
    out[i] = eval_twin(&c, t0, t1);
    out[i].nodes = c.nodes;
  }
  return 0;
}

/* The Python twin, _evaluate_recursive (pure_exact_cpu.py:195-244) with the
 * query set-up of evaluate_external_ccd_query (:247-308): recursive, left
 * half first; contact when the smallest of the three samples (the first of
 * equals, as Python's min) is within eps_space, its time the TOI;
 * separation and undecided report t1.  The motion radius is
 * 0.5 * max(0, t1 - t0) * D with D the point's displacement plus the
 * largest triangle-vertex displacement (PT) or the sum of each edge's
 * largest (EE), computed once.  Returns the status; fills toi, margin,
 * evals (3 per node) and the deepest depth reached. */
typedef struct {
  int status;
  double toi, margin;
  int evals, depth;
} PyOut;

static PyOut eval_py(const Ctx* c, double D, double t0, double t1, int depth) {
  const double mid = 0.5 * (t0 + t1);
  const double st[3] = {t0, mid, t1};
  const double sd[3] = {dist(c, t0), dist(c, mid), dist(c, t1)};
  int k = 0;
  for (int i = 1; i < 3; ++i)
    if (sd[i] < sd[k]) k = i;
  PyOut r = {0, 0.0, 0.0, 3, depth};
  if (sd[k] <= c->eps_space) { r.toi = st[k]; return r; }
  const double lb = sd[1] - 0.5 * smax(0.0, t1 - t0) * D;
  if (lb > c->eps_space) { r.status = 1; r.toi = t1; r.margin = smax(0.0, lb - c->eps_space); return r; }
  if (depth >= c->max_depth || t1 - t0 <= c->eps_time) { r.status = 2; r.toi = t1; return r; }
  const PyOut left = eval_py(c, D, t0, mid, depth + 1);
  r.evals += left.evals;
  if (left.status == 0) { r.toi = left.toi; r.depth = left.depth > depth ? left.depth : depth; return r; }
  const PyOut right = eval_py(c, D, mid, t1, depth + 1);
  r.evals += right.evals;
  r.depth = left.depth > right.depth ? left.depth : right.depth;
  if (right.status == 0) { r.toi = right.toi; return r; }
  r.toi = t1;
  if (left.status == 1 && right.status == 1) { r.status = 1; r.margin = smin(left.margin, right.margin); return r; }
  r.status = 2;
  return r;
}

static double vnorm(const orc_traj* tr) { return v_norm(v_sub(v_load(tr->p1), v_load(tr->p0))); }

/* out: status, toi_upper, safe_margin_lb, nodes = evals / 3, reason_code =
 * deepest depth (the Python result's max_depth) */
int orc_ref_eval_py_batch(int family, const orc_prim* prims, int64_t n, double eps_time, double eps_space,
                          int max_depth, orc_result* out) {
  for (int64_t i = 0; i < n; ++i) {
    if (!isfinite(eps_time) || eps_time <= 0.0 || !isfinite(eps_space) || eps_space <= 0.0 || max_depth < 0)
      return -1;
    const orc_prim* pr = &prims[i];
    double D;
    if (family == 1) {
      double m = vnorm(&pr->v[1]);
      if (m < vnorm(&pr->v[2])) m = vnorm(&pr->v[2]);
      if (m < vnorm(&pr->v[3])) m = vnorm(&pr->v[3]);
      D = vnorm(&pr->v[0]) + m;
    } else {
      D = smax(vnorm(&pr->v[0]), vnorm(&pr->v[1])) + smax(vnorm(&pr->v[2]), vnorm(&pr->v[3]));
    }
    Ctx c = {family, pr, eps_time, eps_space, max_depth, 0};
    g_py_guard = 1;
    const PyOut r = eval_py(&c, D, 0.0, 1.0, 0);
    g_py_guard = 0;
    memset(&out[i], 0, sizeof(out[i]));
    out[i].status = r.status;
    out[i].toi_upper = r.toi;
    out[i].safe_margin_lb = r.margin;
    out[i].nodes = r.evals / 3;
    out[i].reason_code = (uint16_t)r.depth;
  }
  return 0;
}

int orc_ref_eval_batch(int family, const orc_prim* prims, int64_t n, double t0, double t1,
                       double eps_time, double eps_space, int max_depth, orc_result* out) {
  for (int64_t i = 0; i < n; ++i) {
    const int rc = orc_ref_eval(family, &prims[i], t0, t1, eps_time, eps_space, max_depth, &out[i]);
    if (rc != 0) return rc;
  }
  return 0;
}
