// Proxy-scene candidate generation over the gfx950 C-ABI: validation and the
// density statistics on the host (bookkeeping over the scene), raw hits and
// their compaction on the device (p2c_proxy_generate_host /
// p2c_proxy_compact_host).
#include "proxy_hip.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <map>
#include <set>
#include <unordered_map>
#include <utility>

#include "p2c_hip.h"

namespace p2cccd {
namespace {

static_assert(sizeof(RawCandidateHit) == sizeof(p2c_raw_hit), "RawCandidateHit layout");
static_assert(offsetof(RawCandidateHit, motion_bound) == offsetof(p2c_raw_hit, motion_bound),
              "RawCandidateHit layout");
static_assert(sizeof(CandidateRecord) == sizeof(p2c_candidate), "CandidateRecord layout");
static_assert(offsetof(CandidateRecord, candidate_id) == offsetof(p2c_candidate, candidate_id),
              "CandidateRecord layout");
static_assert(offsetof(CandidateRecord, proxy_type_a) == offsetof(p2c_candidate, proxy_type_a),
              "CandidateRecord layout");
static_assert(offsetof(CandidateRecord, rt_hit_count) == offsetof(p2c_candidate, rt_hit_count),
              "CandidateRecord layout");
static_assert(offsetof(CandidateRecord, flags) == offsetof(p2c_candidate, flags), "CandidateRecord layout");

using Clock = std::chrono::steady_clock;

double ElapsedMs(Clock::time_point start) {
  return std::chrono::duration<double, std::milli>(Clock::now() - start).count();
}

double TimeTolerance(double t0, double t1) {
  return 1.0e-12 * std::max({1.0, std::abs(t0), std::abs(t1), std::abs(t1 - t0)});
}

// ValidateGlobalSlabPartition (proxy_scene.cpp:64-96).
Status ValidateSlabPartition(const ProxyScene& scene) {
  std::map<std::uint32_t, std::pair<double, double>> slabs;
  for (const ProxyPrimitive& p : scene.primitives) {
    auto [it, inserted] = slabs.emplace(p.slab_id, std::make_pair(p.t0, p.t1));
    if (!inserted) {
      const double tol = TimeTolerance(it->second.first, it->second.second);
      if (std::abs(it->second.first - p.t0) > tol || std::abs(it->second.second - p.t1) > tol)
        return Status::Error("proxy scene slab_id must map to one global time interval");
    }
  }
  std::vector<std::pair<std::uint32_t, std::pair<double, double>>> ordered(slabs.begin(), slabs.end());
  std::sort(ordered.begin(), ordered.end(), [](const auto& a, const auto& b) {
    if (a.second.first == b.second.first) return a.first < b.first;
    return a.second.first < b.second.first;
  });
  for (std::size_t i = 1; i < ordered.size(); ++i) {
    const auto& a = ordered[i - 1].second;
    const auto& b = ordered[i].second;
    const double tol = std::max(TimeTolerance(a.first, a.second), TimeTolerance(b.first, b.second));
    if (a.first < b.second - tol && b.first < a.second - tol)
      return Status::Error("proxy scene global slab intervals must not overlap");
  }
  return Status::Ok();
}

// Host arrays in the C-ABI's proxy-set layout.
struct PackedSet {
  std::vector<std::uint32_t> object_id, patch_id, slab_id, proxy_type;
  std::vector<double> t0, t1, bounds, motion;
  p2c_proxy_set set{};

  explicit PackedSet(const ProxyScene& scene) {
    const std::size_t n = scene.primitives.size();
    object_id.resize(n);
    patch_id.resize(n);
    slab_id.resize(n);
    proxy_type.resize(n);
    t0.resize(n);
    t1.resize(n);
    bounds.resize(6 * n);
    motion.resize(4 * n);
    for (std::size_t i = 0; i < n; ++i) {
      const ProxyPrimitive& p = scene.primitives[i];
      object_id[i] = p.object_id;
      patch_id[i] = p.patch_id;
      slab_id[i] = p.slab_id;
      proxy_type[i] = static_cast<std::uint32_t>(p.proxy_type);
      t0[i] = p.t0;
      t1[i] = p.t1;
      for (int k = 0; k < 3; ++k) {
        bounds[6 * i + k] = p.bounds.min[k];
        bounds[6 * i + 3 + k] = p.bounds.max[k];
      }
      // MergeMotionBound's four fields (candidate_buffer.cpp:37-46)
      motion[4 * i + 0] = p.motion_bound.translation_bound;
      motion[4 * i + 1] = p.motion_bound.rotation_angle;
      motion[4 * i + 2] = p.motion_bound.radial_motion_bound;
      motion[4 * i + 3] = p.motion_bound.conservative_radius;
    }
    set.object_id = object_id.data();
    set.patch_id = patch_id.data();
    set.slab_id = slab_id.data();
    set.proxy_type = proxy_type.data();
    set.t0 = t0.data();
    set.t1 = t1.data();
    set.bounds = bounds.data();
    set.motion = motion.data();
    set.n = static_cast<std::int64_t>(n);
    set.query_id = scene.query_id;
  }
};

Status FromRc(int rc) {
  if (rc == P2C_OK) return Status::Ok();
  return Status::Error(p2c_last_error());
}

template <typename T>
struct HostArray {
  T* p = nullptr;
  ~HostArray() { p2c_host_free(p); }
};

}  // namespace

Status ValidateProxyScene(const ProxyScene& scene) {
  if (scene.query_id == 0) return Status::Error("ProxyScene.query_id is required");
  for (std::size_t i = 0; i < scene.primitives.size(); ++i) {
    const ProxyPrimitive& p = scene.primitives[i];
    if (p.proxy_id != i) return Status::Error("proxy_id must match primitive array index");
    if (p.proxy_type != ProxyType::kSweptAabb && p.proxy_type != ProxyType::kCapsule)
      return Status::Error("proxy primitive has an unsupported proxy type");
    if (!std::isfinite(p.t0) || !std::isfinite(p.t1) || p.t0 >= p.t1)
      return Status::Error("proxy primitive must satisfy finite t0 < t1");
    for (int k = 0; k < 3; ++k) {
      if (!std::isfinite(p.bounds.min[k]) || !std::isfinite(p.bounds.max[k]))
        return Status::Error("proxy bounds contain a non-finite value");
      if (p.bounds.min[k] > p.bounds.max[k]) return Status::Error("proxy bounds min cannot exceed max");
    }
  }
  return ValidateSlabPartition(scene);
}

Status ComputeCandidateDensityStats(const ProxyScene& scene, const RawCandidateBuffer& raw_buffer,
                                    const std::vector<CandidateRecord>& candidates, const RtCandidateTiming& timing,
                                    const std::string& backend_name, CandidateDensityStats* stats) {
  if (stats == nullptr) return Status::Error("candidate density stats output pointer is null");
  if (auto s = ValidateProxyScene(scene); !s.ok) return s;
  std::set<std::uint32_t> objects, slabs;
  // cross-object same-slab pairs: per slab, all pairs minus same-object pairs
  std::map<std::uint32_t, std::pair<std::uint64_t, std::map<std::uint32_t, std::uint64_t>>> per_slab;
  for (const ProxyPrimitive& p : scene.primitives) {
    objects.insert(p.object_id);
    slabs.insert(p.slab_id);
    auto& e = per_slab[p.slab_id];
    ++e.first;
    ++e.second[p.object_id];
  }
  std::uint64_t cross = 0;
  for (const auto& [slab, e] : per_slab) {
    std::uint64_t same = 0;
    for (const auto& [obj, k] : e.second) same += k * (k - 1) / 2;
    cross += e.first * (e.first - 1) / 2 - same;
  }
  auto ratio = [](std::uint64_t a, std::uint64_t b) { return b == 0 ? 0.0 : double(a) / double(b); };
  std::uint64_t rt_sum = 0;
  for (const CandidateRecord& c : candidates) rt_sum += c.rt_hit_count;
  CandidateDensityStats d;
  d.query_id = scene.query_id;
  d.proxy_count = scene.primitives.size();
  d.object_count = objects.size();
  d.slab_count = slabs.size();
  d.cross_object_same_slab_pair_count = cross;
  d.raw_hit_count = raw_buffer.hits.size();
  d.compact_candidate_count = candidates.size();
  d.raw_hits_per_proxy = ratio(d.raw_hit_count, d.proxy_count);
  d.candidates_per_proxy = ratio(d.compact_candidate_count, d.proxy_count);
  d.candidates_per_slab = ratio(d.compact_candidate_count, d.slab_count);
  d.aabb_overlap_ratio = ratio(d.raw_hit_count, d.cross_object_same_slab_pair_count);
  d.avg_rt_hits_per_candidate = ratio(rt_sum, d.compact_candidate_count);
  d.timing = timing;
  d.backend_name = backend_name;
  *stats = std::move(d);
  return Status::Ok();
}

Status CompactRawCandidatesHip(const ProxyScene& scene, const RawCandidateBuffer& raw_buffer,
                               std::vector<CandidateRecord>* candidates) {
  if (candidates == nullptr) return Status::Error("candidate output pointer is null");
  if (auto s = ValidateProxyScene(scene); !s.ok) return s;
  PackedSet packed(scene);
  HostArray<p2c_candidate> cand;
  std::int64_t n_cand = 0;
  const auto* raw = reinterpret_cast<const p2c_raw_hit*>(raw_buffer.hits.data());
  if (auto s = FromRc(p2c_proxy_compact_host(&packed.set, raw, static_cast<std::int64_t>(raw_buffer.hits.size()),
                                             &cand.p, &n_cand));
      !s.ok)
    return s;
  candidates->resize(static_cast<std::size_t>(n_cand));
  if (n_cand > 0) std::memcpy(candidates->data(), cand.p, sizeof(CandidateRecord) * std::size_t(n_cand));
  return Status::Ok();
}

Status GenerateCandidatesHip(const ProxyScene& scene, std::uint64_t query_id, CandidateGenerationResult* result) {
  if (result == nullptr) return Status::Error("candidate generation result output pointer is null");
  if (query_id == 0) return Status::Error("query_id is required");
  if (query_id != scene.query_id) return Status::Error("query_id must match ProxyScene.query_id");
  CandidateGenerationResult g;
  g.backend_name = kHipCandidateBackend;
  const Clock::time_point total_start = Clock::now();
  Clock::time_point stage = Clock::now();
  if (auto s = ValidateProxyScene(scene); !s.ok) return s;
  g.timing.build_ms = ElapsedMs(stage);
  g.timing.update_ms = 0.0;
  PackedSet packed(scene);
  HostArray<p2c_raw_hit> raw;
  HostArray<p2c_candidate> cand;
  std::int64_t n_raw = 0, n_cand = 0;
  double stage_ms[2] = {0.0, 0.0};
  if (auto s = FromRc(p2c_proxy_generate_host(&packed.set, &raw.p, &n_raw, &cand.p, &n_cand, stage_ms)); !s.ok)
    return s;
  g.timing.trace_ms = stage_ms[0];    // device: count, scan, fill
  g.timing.compact_ms = stage_ms[1];  // device: validation, key sort, merge
  g.raw_buffer.hits.resize(static_cast<std::size_t>(n_raw));
  if (n_raw > 0) std::memcpy(g.raw_buffer.hits.data(), raw.p, sizeof(RawCandidateHit) * std::size_t(n_raw));
  g.candidates.resize(static_cast<std::size_t>(n_cand));
  if (n_cand > 0) std::memcpy(g.candidates.data(), cand.p, sizeof(CandidateRecord) * std::size_t(n_cand));
  stage = Clock::now();
  g.timing.total_ms = ElapsedMs(total_start);
  if (auto s = ComputeCandidateDensityStats(scene, g.raw_buffer, g.candidates, g.timing, g.backend_name, &g.density);
      !s.ok)
    return s;
  g.timing.stats_ms = ElapsedMs(stage);
  g.timing.total_ms = ElapsedMs(total_start);
  g.density.timing = g.timing;
  *result = std::move(g);
  return Status::Ok();
}

}  // namespace p2cccd
