// Rigid-patch proxies and proxy-scene construction on the host: the swept
// bounds of patches under rigid motion (translation lerp, quaternion slerp)
// that the proxy-scene candidate generator consumes.  The arithmetic follows
// src/cpp/geometry/motion_utils.cpp:196-325, src/cpp/geometry/proxy.cpp:36-133
// and src/cpp/rt_candidate/proxy_scene.cpp:98-213 operation for operation
// (same products, sums, sqrt/acos/sin/cos calls in the same order), so every
// bound and motion bound equals the reference's bit for bit; validation
// messages are the reference's texts.  This is scene set-up, run once per
// scene, not the per-step path.
#include <algorithm>
#include <cmath>
#include <string>

#include "proxy_hip.h"

namespace p2cccd {
namespace {

using Vec3 = std::array<double, 3>;
using Quat = std::array<double, 4>;

bool Finite3(const Vec3& v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }
bool Finite4(const Quat& q) { return Finite3({q[0], q[1], q[2]}) && std::isfinite(q[3]); }
double Dot3(const Vec3& a, const Vec3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
double Dot4(const Quat& a, const Quat& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3]; }
Vec3 Sum3(const Vec3& a, const Vec3& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
Vec3 Times3(const Vec3& a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
double Length3(const Vec3& a) { return std::sqrt(Dot3(a, a)); }

Status UnitQuaternion(const Quat& in, Quat* out) {
  if (!Finite4(in)) return Status::Error("quaternion contains a non-finite value");
  const double norm = std::sqrt(Dot4(in, in));
  if (!std::isfinite(norm) || norm <= 0.0) return Status::Error("quaternion norm must be positive and finite");
  *out = {in[0] / norm, in[1] / norm, in[2] / norm, in[3] / norm};
  return Status::Ok();
}

Status CheckPose(const PoseSample& pose, const char* name) {
  if (!Finite3(pose.translation)) return Status::Error(std::string(name) + ".translation contains a non-finite value");
  Quat unit{};
  return UnitQuaternion(pose.rotation_xyzw, &unit);
}

// Spherical interpolation of unit quaternions along the shorter arc; nearly
// parallel ones (|cos| > 0.9995) are interpolated linearly and renormalised.
Quat Slerp(const Quat& q0, Quat q1, double alpha) {
  double c = Dot4(q0, q1);
  if (c < 0.0) {
    q1 = {-q1[0], -q1[1], -q1[2], -q1[3]};
    c = -c;
  }
  c = std::clamp(c, -1.0, 1.0);
  if (c > 0.9995) {
    const Quat q{(1.0 - alpha) * q0[0] + alpha * q1[0], (1.0 - alpha) * q0[1] + alpha * q1[1],
                 (1.0 - alpha) * q0[2] + alpha * q1[2], (1.0 - alpha) * q0[3] + alpha * q1[3]};
    const double norm = std::sqrt(Dot4(q, q));
    return {q[0] / norm, q[1] / norm, q[2] / norm, q[3] / norm};
  }
  const double full = std::acos(c);
  const double part = full * alpha;
  const double s_part = std::sin(part);
  const double s_full = std::sin(full);
  const double w0 = std::cos(part) - c * s_part / s_full;
  const double w1 = s_part / s_full;
  return {w0 * q0[0] + w1 * q1[0], w0 * q0[1] + w1 * q1[1], w0 * q0[2] + w1 * q1[2], w0 * q0[3] + w1 * q1[3]};
}

// q v q* for a unit q = (u, s): 2 (u.v) u + (s^2 - u.u) v + 2 s (u x v).
Vec3 Rotate(const Quat& q, const Vec3& v) {
  const Vec3 u{q[0], q[1], q[2]};
  const double s = q[3];
  const Vec3 along = Times3(u, 2.0 * Dot3(u, v));
  const Vec3 keep = Times3(v, s * s - Dot3(u, u));
  const Vec3 cross{u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
  return Sum3(Sum3(along, keep), Times3(cross, 2.0 * s));
}

double SegmentTolerance(const MotionSegment& m) { return 1.0e-12 * std::max(1.0, m.t1 - m.t0); }

Status CheckSlab(const MotionSegment& m, const TimeSlab& slab) {
  if (!std::isfinite(slab.t0) || !std::isfinite(slab.t1) || slab.t0 >= slab.t1)
    return Status::Error("time slab must satisfy finite t0 < t1");
  const double tol = SegmentTolerance(m);
  if (slab.t0 < m.t0 - tol || slab.t1 > m.t1 + tol) return Status::Error("time slab must lie inside the motion segment");
  return Status::Ok();
}

Status CheckEps(double eps) {
  if (!std::isfinite(eps) || eps < 0.0) return Status::Error("eps_proxy must be finite and non-negative");
  return Status::Ok();
}

}  // namespace

Status ValidateMotionSegment(const MotionSegment& segment) {
  if (!std::isfinite(segment.t0) || !std::isfinite(segment.t1) || segment.t0 >= segment.t1)
    return Status::Error("motion segment must satisfy finite t0 < t1");
  if (auto s = CheckPose(segment.pose_t0, "pose_t0"); !s.ok) return s;
  return CheckPose(segment.pose_t1, "pose_t1");
}

Status GenerateUniformTimeSlabs(double t0, double t1, std::uint32_t slab_count, std::vector<TimeSlab>* slabs) {
  if (slabs == nullptr) return Status::Error("slab output pointer is null");
  if (!std::isfinite(t0) || !std::isfinite(t1) || t0 >= t1)
    return Status::Error("slab interval must satisfy finite t0 < t1");
  if (slab_count == 0) return Status::Error("slab_count must be positive");
  slabs->clear();
  const double dt = (t1 - t0) / static_cast<double>(slab_count);
  for (std::uint32_t k = 0; k < slab_count; ++k)
    slabs->push_back(TimeSlab{k, k == 0 ? t0 : t0 + dt * k, k + 1 == slab_count ? t1 : t0 + dt * (k + 1)});
  return Status::Ok();
}

Status InterpolateRigidMotion(const MotionSegment& segment, double t, PoseSample* pose) {
  if (pose == nullptr) return Status::Error("pose output pointer is null");
  if (auto s = ValidateMotionSegment(segment); !s.ok) return s;
  const double tol = SegmentTolerance(segment);
  if (!std::isfinite(t) || t < segment.t0 - tol || t > segment.t1 + tol)
    return Status::Error("interpolation time must lie inside the motion segment");
  const double alpha = (std::clamp(t, segment.t0, segment.t1) - segment.t0) / (segment.t1 - segment.t0);
  const Vec3& a = segment.pose_t0.translation;
  const Vec3& b = segment.pose_t1.translation;
  pose->translation = {(1.0 - alpha) * a[0] + alpha * b[0], (1.0 - alpha) * a[1] + alpha * b[1],
                       (1.0 - alpha) * a[2] + alpha * b[2]};
  Quat q0{}, q1{};
  if (auto s = UnitQuaternion(segment.pose_t0.rotation_xyzw, &q0); !s.ok) return s;
  if (auto s = UnitQuaternion(segment.pose_t1.rotation_xyzw, &q1); !s.ok) return s;
  pose->rotation_xyzw = Slerp(q0, q1, alpha);
  return Status::Ok();
}

Status TransformPoint(const PoseSample& pose, const std::array<double, 3>& local_point,
                      std::array<double, 3>* world_point) {
  if (world_point == nullptr) return Status::Error("world point output pointer is null");
  if (!Finite3(local_point)) return Status::Error("local point contains a non-finite value");
  if (auto s = CheckPose(pose, "pose"); !s.ok) return s;
  Quat q{};
  if (auto s = UnitQuaternion(pose.rotation_xyzw, &q); !s.ok) return s;
  *world_point = Sum3(Rotate(q, local_point), pose.translation);
  return Status::Ok();
}

Status ComputeRotationAngle(const PoseSample& pose0, const PoseSample& pose1, double* angle) {
  if (angle == nullptr) return Status::Error("angle output pointer is null");
  Quat q0{}, q1{};
  if (auto s = UnitQuaternion(pose0.rotation_xyzw, &q0); !s.ok) return s;
  if (auto s = UnitQuaternion(pose1.rotation_xyzw, &q1); !s.ok) return s;
  *angle = 2.0 * std::acos(std::clamp(std::abs(Dot4(q0, q1)), 0.0, 1.0));
  return Status::Ok();
}

Status ComputePatchMotionBound(const Patch& patch, const MotionSegment& segment, const TimeSlab& slab,
                               PatchMotionBound* bound) {
  if (bound == nullptr) return Status::Error("motion bound output pointer is null");
  if (!Finite3(patch.local_center)) return Status::Error("patch.local_center contains a non-finite value");
  if (!std::isfinite(patch.radius) || patch.radius < 0.0)
    return Status::Error("patch.radius must be finite and non-negative");
  if (auto s = ValidateMotionSegment(segment); !s.ok) return s;
  if (auto s = CheckSlab(segment, slab); !s.ok) return s;
  PoseSample p0, p1;
  if (auto s = InterpolateRigidMotion(segment, slab.t0, &p0); !s.ok) return s;
  if (auto s = InterpolateRigidMotion(segment, slab.t1, &p1); !s.ok) return s;
  PatchMotionBound r;
  r.patch_id = patch.patch_id;
  r.t0 = slab.t0;
  r.t1 = slab.t1;
  if (auto s = TransformPoint(p0, patch.local_center, &r.center_t0); !s.ok) return s;
  if (auto s = TransformPoint(p1, patch.local_center, &r.center_t1); !s.ok) return s;
  if (auto s = ComputeRotationAngle(p0, p1, &r.rotation_angle); !s.ok) return s;
  const Vec3 d{p0.translation[0] - p1.translation[0], p0.translation[1] - p1.translation[1],
               p0.translation[2] - p1.translation[2]};
  r.translation_bound = Length3(d);
  // a point at distance rho from the rotation centre moves by at most the
  // chord 2 rho sin(angle / 2)
  const double chord = std::sin(0.5 * r.rotation_angle);
  r.center_rotation_bound = 2.0 * Length3(patch.local_center) * chord;
  r.surface_rotation_bound = 2.0 * patch.radius * chord;
  r.radial_motion_bound = r.translation_bound + r.center_rotation_bound + r.surface_rotation_bound;
  r.conservative_radius = patch.radius + r.center_rotation_bound;
  *bound = r;
  return Status::Ok();
}

Status BuildEndpointSweptAabbProxy(const Patch& patch, const MotionSegment& segment, const TimeSlab& slab,
                                   double eps_proxy, EndpointSweptAabbProxy* proxy) {
  if (proxy == nullptr) return Status::Error("endpoint swept AABB proxy output pointer is null");
  if (auto s = CheckEps(eps_proxy); !s.ok) return s;
  PatchMotionBound b;
  if (auto s = ComputePatchMotionBound(patch, segment, slab, &b); !s.ok) return s;
  EndpointSweptAabbProxy out;
  out.patch_id = patch.patch_id;
  out.motion_bound = b;
  for (int a = 0; a < 3; ++a) {
    out.bounds.min[a] = std::min(b.center_t0[a], b.center_t1[a]) - b.conservative_radius;
    out.bounds.max[a] = std::max(b.center_t0[a], b.center_t1[a]) + b.conservative_radius;
  }
  for (int a = 0; a < 3; ++a) {  // InflateAabb
    if (!std::isfinite(out.bounds.min[a]) || !std::isfinite(out.bounds.max[a]))
      return Status::Error("AABB contains a non-finite bound");
    if (out.bounds.min[a] > out.bounds.max[a]) return Status::Error("AABB min cannot exceed max");
  }
  for (int a = 0; a < 3; ++a) {
    out.bounds.min[a] -= eps_proxy;
    out.bounds.max[a] += eps_proxy;
  }
  *proxy = out;
  return Status::Ok();
}

Status BuildCapsuleProxy(const Patch& patch, const MotionSegment& segment, const TimeSlab& slab, double eps_proxy,
                         CapsuleProxy* proxy) {
  if (proxy == nullptr) return Status::Error("capsule proxy output pointer is null");
  if (auto s = CheckEps(eps_proxy); !s.ok) return s;
  PatchMotionBound b;
  if (auto s = ComputePatchMotionBound(patch, segment, slab, &b); !s.ok) return s;
  CapsuleProxy out;
  out.patch_id = patch.patch_id;
  out.endpoint0 = b.center_t0;
  out.endpoint1 = b.center_t1;
  out.radius = b.conservative_radius;
  out.motion_bound = b;
  if (!std::isfinite(out.radius) || out.radius < 0.0)  // InflateCapsule
    return Status::Error("capsule radius must be finite and non-negative");
  out.radius += eps_proxy;
  *proxy = out;
  return Status::Ok();
}

Status BuildProxyScene(const ProxySceneBuildInput& input, ProxyScene* scene) {
  if (scene == nullptr) return Status::Error("proxy scene output pointer is null");
  if (input.query_id == 0) return Status::Error("ProxySceneBuildInput.query_id is required");
  if (input.objects.empty()) return Status::Error("ProxySceneBuildInput.objects is empty");
  ProxyScene built;
  built.query_id = input.query_id;
  for (const ProxyObjectBuildInput& obj : input.objects) {
    if (obj.proxy_type != ProxyType::kSweptAabb && obj.proxy_type != ProxyType::kCapsule)
      return Status::Error("ProxyObjectBuildInput.proxy_type must be concrete");
    if (obj.patches.empty()) return Status::Error("ProxyObjectBuildInput.patches is empty");
    if (obj.motion_segments.empty()) return Status::Error("ProxyObjectBuildInput.motion_segments is empty");
    if (obj.slabs_per_motion_segment == 0) return Status::Error("slabs_per_motion_segment must be positive");
    if (!std::isfinite(obj.eps_proxy) || obj.eps_proxy < 0.0)
      return Status::Error("eps_proxy must be finite and non-negative");
    for (std::uint32_t seg = 0; seg < obj.motion_segments.size(); ++seg) {
      const MotionSegment& m = obj.motion_segments[seg];
      if (auto s = ValidateMotionSegment(m); !s.ok) return s;
      std::vector<TimeSlab> slabs;
      if (auto s = GenerateUniformTimeSlabs(m.t0, m.t1, obj.slabs_per_motion_segment, &slabs); !s.ok) return s;
      for (TimeSlab slab : slabs) {
        slab.slab_id = seg * obj.slabs_per_motion_segment + slab.slab_id;  // one global slab grid per scene
        for (const Patch& patch : obj.patches) {
          ProxyPrimitive p;
          p.proxy_id = static_cast<std::uint32_t>(built.primitives.size());
          p.object_id = obj.object_id;
          p.patch_id = patch.patch_id;
          p.slab_id = slab.slab_id;
          p.motion_segment_id = seg;
          p.proxy_type = obj.proxy_type;
          p.t0 = slab.t0;
          p.t1 = slab.t1;
          if (obj.proxy_type == ProxyType::kSweptAabb) {
            EndpointSweptAabbProxy px;
            if (auto s = BuildEndpointSweptAabbProxy(patch, m, slab, obj.eps_proxy, &px); !s.ok) return s;
            p.bounds = px.bounds;
            p.motion_bound = px.motion_bound;
          } else {
            CapsuleProxy cx;
            if (auto s = BuildCapsuleProxy(patch, m, slab, obj.eps_proxy, &cx); !s.ok) return s;
            p.capsule = cx;
            for (int a = 0; a < 3; ++a) {
              p.bounds.min[a] = std::min(cx.endpoint0[a], cx.endpoint1[a]) - cx.radius;
              p.bounds.max[a] = std::max(cx.endpoint0[a], cx.endpoint1[a]) + cx.radius;
            }
            p.motion_bound = cx.motion_bound;
          }
          for (int a = 0; a < 3; ++a) {
            if (!std::isfinite(p.bounds.min[a]) || !std::isfinite(p.bounds.max[a]))
              return Status::Error("proxy bounds contain a non-finite value");
            if (p.bounds.min[a] > p.bounds.max[a]) return Status::Error("proxy bounds min cannot exceed max");
          }
          built.primitives.push_back(std::move(p));
        }
      }
    }
  }
  if (auto s = ValidateProxyScene(built); !s.ok) return s;
  *scene = std::move(built);
  return Status::Ok();
}

}  // namespace p2cccd
