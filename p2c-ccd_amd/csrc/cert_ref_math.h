// Reference-parity interval certificate: FP64 distance and motion-bound math
// shared by the gfx950 kernel and the library's host CPU entry point.
//
// Every expression keeps the operand order of the reference
// (src/cpp/certificate/certificate_engine.cpp:27-222) and the translation
// units are compiled with -ffp-contract=off, so device and host results are
// bit-identical to the reference CPU build (no FMA contraction; IEEE
// correctly rounded sqrt and division on both sides).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "p2c_hip.h"

#define P2C_HD __host__ __device__ inline

namespace p2c {

struct D3 {
  double x, y, z;
};

P2C_HD D3 add3(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
P2C_HD D3 sub3(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
P2C_HD D3 mul3(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
P2C_HD double dot3(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
P2C_HD D3 cross3(D3 a, D3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
P2C_HD double norm3(D3 a) { return sqrt(dot3(a, a)); }

// std::max / std::min / std::clamp, including their NaN behaviour.
P2C_HD double std_max(double a, double b) { return (a < b) ? b : a; }
P2C_HD double std_min(double a, double b) { return (b < a) ? b : a; }
P2C_HD double std_clamp(double v, double lo, double hi) {
  return (v < lo) ? lo : ((hi < v) ? hi : v);
}

constexpr double kRefNumEps = 1.0e-14;  // certificate_engine.cpp:20

// One primitive's 4 trajectories as start/end positions.
struct RefPrim {
  D3 p0[4];
  D3 p1[4];
};

P2C_HD RefPrim load_prim(const p2c_prim& p) {
  RefPrim r;
  for (int i = 0; i < 4; ++i) {
    r.p0[i] = {p.v[i].position_t0[0], p.v[i].position_t0[1], p.v[i].position_t0[2]};
    r.p1[i] = {p.v[i].position_t1[0], p.v[i].position_t1[1], p.v[i].position_t1[2]};
  }
  return r;
}

// (1 - t) * p0 + t * p1, evaluated per component in that order.
P2C_HD D3 lerp_at(D3 p0, D3 p1, double t) { return add3(mul3(p0, 1.0 - t), mul3(p1, t)); }

P2C_HD double dist_point_point(D3 a, D3 b) { return norm3(sub3(a, b)); }

P2C_HD double dist_point_segment(D3 p, D3 a, D3 b) {
  const D3 ab = sub3(b, a);
  const double ab2 = dot3(ab, ab);
  if (ab2 <= kRefNumEps) return dist_point_point(p, a);
  const double s = std_clamp(dot3(sub3(p, a), ab) / ab2, 0.0, 1.0);
  return dist_point_point(p, add3(a, mul3(ab, s)));
}

// Clamped segment-segment closest distance with the reference's `denom != 0`
// guard and re-clamping order; `py` takes the Python twin's guard
// `abs(denom) > 1e-14` (pure_exact_cpu.py:126) instead.
P2C_HD double dist_segment_segment(D3 p1, D3 q1, D3 p2, D3 q2, bool py = false) {
  const D3 d1 = sub3(q1, p1);
  const D3 d2 = sub3(q2, p2);
  const D3 r = sub3(p1, p2);
  const double a = dot3(d1, d1);
  const double e = dot3(d2, d2);
  const double f = dot3(d2, r);
  double s = 0.0, t = 0.0;
  if (a <= kRefNumEps && e <= kRefNumEps) return dist_point_point(p1, p2);
  if (a <= kRefNumEps) {
    t = std_clamp(f / e, 0.0, 1.0);
  } else {
    const double c = dot3(d1, r);
    if (e <= kRefNumEps) {
      s = std_clamp(-c / a, 0.0, 1.0);
    } else {
      const double b = dot3(d1, d2);
      const double denom = a * e - b * b;
      if (py ? (denom < 0.0 ? -denom : denom) > 1.0e-14 : denom != 0.0) s = std_clamp((b * f - c * e) / denom, 0.0, 1.0);
      t = (b * s + f) / e;
      if (t < 0.0) {
        t = 0.0;
        s = std_clamp(-c / a, 0.0, 1.0);
      } else if (t > 1.0) {
        t = 1.0;
        s = std_clamp((b - c) / a, 0.0, 1.0);
      }
    }
  }
  return dist_point_point(add3(p1, mul3(d1, s)), add3(p2, mul3(d2, t)));
}

// Voronoi-region point-triangle distance; a triangle whose |n|^2 <= 1e-14 is
// treated as its three edges (first smallest kept, as std::min({...})).
P2C_HD double dist_point_triangle(D3 p, D3 a, D3 b, D3 c) {
  const D3 ab = sub3(b, a);
  const D3 ac = sub3(c, a);
  const D3 n = cross3(ab, ac);
  if (dot3(n, n) <= kRefNumEps) {
    double m = dist_point_segment(p, a, b);
    const double y = dist_point_segment(p, b, c);
    const double z = dist_point_segment(p, c, a);
    if (y < m) m = y;
    if (z < m) m = z;
    return m;
  }
  const D3 ap = sub3(p, a);
  const double d1 = dot3(ab, ap);
  const double d2 = dot3(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return dist_point_point(p, a);
  const D3 bp = sub3(p, b);
  const double d3 = dot3(ab, bp);
  const double d4 = dot3(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) return dist_point_point(p, b);
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
    const double w = d1 / (d1 - d3);
    return dist_point_point(p, add3(a, mul3(ab, w)));
  }
  const D3 cp = sub3(p, c);
  const double d5 = dot3(ab, cp);
  const double d6 = dot3(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return dist_point_point(p, c);
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
    const double w = d2 / (d2 - d6);
    return dist_point_point(p, add3(a, mul3(ac, w)));
  }
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    return dist_point_point(p, add3(b, mul3(sub3(c, b), w)));
  }
  return fabs(dot3(ap, n) / norm3(n));
}

template <int kFamily>
P2C_HD double ref_distance(const RefPrim& q, double t, bool py = false) {
  const D3 x0 = lerp_at(q.p0[0], q.p1[0], t);
  const D3 x1 = lerp_at(q.p0[1], q.p1[1], t);
  const D3 x2 = lerp_at(q.p0[2], q.p1[2], t);
  const D3 x3 = lerp_at(q.p0[3], q.p1[3], t);
  if (kFamily == P2C_FAMILY_POINT_TRIANGLE) return dist_point_triangle(x0, x1, x2, x3);
  return dist_segment_segment(x0, x1, x2, x3, py);
}

// 0.5 * max(0, t1 - t0) * |p1 - p0|, per vertex.
P2C_HD double max_disp(D3 p0, D3 p1, double t0, double t1) {
  const double len = std_max(0.0, t1 - t0);
  return 0.5 * len * norm3(sub3(p1, p0));
}

// Lipschitz motion radius of the interval: point + max(triangle vertices) for
// PT, max(edge a) + max(edge b) for EE.
template <int kFamily>
P2C_HD double ref_motion_radius(const RefPrim& q, double t0, double t1) {
  if (kFamily == P2C_FAMILY_POINT_TRIANGLE) {
    const double pd = max_disp(q.p0[0], q.p1[0], t0, t1);
    double td = max_disp(q.p0[1], q.p1[1], t0, t1);
    const double y = max_disp(q.p0[2], q.p1[2], t0, t1);
    const double z = max_disp(q.p0[3], q.p1[3], t0, t1);
    if (td < y) td = y;
    if (td < z) td = z;
    return pd + td;
  }
  const double ea = std_max(max_disp(q.p0[0], q.p1[0], t0, t1), max_disp(q.p0[1], q.p1[1], t0, t1));
  const double eb = std_max(max_disp(q.p0[2], q.p1[2], t0, t1), max_disp(q.p0[3], q.p1[3], t0, t1));
  return ea + eb;
}

// The same radius with each vertex's |p1 - p0| computed once per primitive:
// max_disp is (0.5 * len) * |p1 - p0|, so reusing the norm gives the same
// product bit for bit.
P2C_HD double disp_n(double half_len, double vn) { return half_len * vn; }

template <int kFamily>
P2C_HD double ref_motion_radius_n(const double vn[4], double t0, double t1) {
  const double h = 0.5 * std_max(0.0, t1 - t0);
  if (kFamily == P2C_FAMILY_POINT_TRIANGLE) {
    const double pd = disp_n(h, vn[0]);
    double td = disp_n(h, vn[1]);
    const double y = disp_n(h, vn[2]);
    const double z = disp_n(h, vn[3]);
    if (td < y) td = y;
    if (td < z) td = z;
    return pd + td;
  }
  const double ea = std_max(disp_n(h, vn[0]), disp_n(h, vn[1]));
  const double eb = std_max(disp_n(h, vn[2]), disp_n(h, vn[3]));
  return ea + eb;
}

P2C_HD bool prim_valid(const p2c_prim& p) {
  if (p.id_a < 0 || p.id_b < 0) return false;
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 3; ++k)
      if (!__builtin_isfinite(p.v[i].position_t0[k]) || !__builtin_isfinite(p.v[i].position_t1[k])) return false;
  return true;
}

P2C_HD void init_result(p2c_prim_result* r, int status, uint32_t family, double t0, double t1) {
  r->status = (uint8_t)status;
  r->witness_family = 0;
  r->reason_code = 0;
  r->covered_feature_mask = status == P2C_STATUS_SEPARATION ? family : 0u;
  r->interval_t0 = t0;
  r->interval_t1 = t1;
  r->toi_upper = t1;
  r->safe_margin_lb = 0.0;
  r->witness_id_a = -1;
  r->witness_id_b = -1;
  r->next_refinement_mode = 0;
  r->pad[0] = r->pad[1] = r->pad[2] = 0;
  r->nodes = 0;
}

}  // namespace p2c
