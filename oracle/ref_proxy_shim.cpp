// Oracle (test infrastructure only): a C-ABI shim over the reference's own
// proxy-scene candidate code, compiled from the reference sources where they
// lie (src/cpp/rt_candidate/{candidate_buffer,proxy_scene,candidate_stats}.cpp,
// src/cpp/geometry/{proxy,motion_utils}.cpp, src/cpp/common/validators.cpp)
// into oracle/_ref/libp2c_ref.so by oracle/Makefile.  Used in this container
// to write golden vectors (tests/golden/make_proxy_golden.py) and to check the
// C restatement in oracle/proxy.c.  Never part of the shipped library.
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle.h"
#include "rt_candidate/candidate_buffer.h"
#include "rt_candidate/candidate_stats.h"
#include "rt_candidate/proxy_scene.h"

namespace {

p2cccd::ProxyScene MakeScene(uint64_t query_id, const orc_proxy_set* s) {
  p2cccd::ProxyScene scene;
  scene.query_id = query_id;
  scene.primitives.resize(static_cast<size_t>(s->n));
  for (int64_t i = 0; i < s->n; ++i) {
    auto& p = scene.primitives[static_cast<size_t>(i)];
    p.proxy_id = static_cast<uint32_t>(i);
    p.object_id = s->object_id[i];
    p.patch_id = s->patch_id[i];
    p.slab_id = s->slab_id[i];
    p.proxy_type = static_cast<p2cccd::ProxyType>(s->proxy_type[i]);
    p.t0 = s->t0[i];
    p.t1 = s->t1[i];
    for (int k = 0; k < 3; ++k) {
      p.bounds.min[k] = s->bounds[6 * i + k];
      p.bounds.max[k] = s->bounds[6 * i + 3 + k];
    }
    p.motion_bound.translation_bound = s->motion[4 * i + 0];
    p.motion_bound.rotation_angle = s->motion[4 * i + 1];
    p.motion_bound.radial_motion_bound = s->motion[4 * i + 2];
    p.motion_bound.conservative_radius = s->motion[4 * i + 3];
  }
  return scene;
}

void CopyError(const p2cccd::Status& st, char* err, int errlen) {
  if (err != nullptr && errlen > 0) {
    std::strncpy(err, st.message.c_str(), static_cast<size_t>(errlen - 1));
    err[errlen - 1] = 0;
  }
}

}  // namespace

extern "C" int ref_proxy_generate(uint64_t query_id, const orc_proxy_set* set, int64_t raw_cap, orc_raw_hit* raw,
                                  int64_t* n_raw, int64_t cand_cap, orc_candidate* cand, int64_t* n_cand,
                                  orc_density* density, char* err, int errlen) {
  const p2cccd::ProxyScene scene = MakeScene(query_id, set);
  p2cccd::RawCandidateBuffer buffer;
  if (auto st = p2cccd::GenerateRawCandidatesCpu(scene, query_id, &buffer); !st.ok) {
    CopyError(st, err, errlen);
    return -1;
  }
  std::vector<p2cccd::CandidateRecord> records;
  if (auto st = p2cccd::CompactRawCandidates(scene, buffer, &records); !st.ok) {
    CopyError(st, err, errlen);
    return -2;
  }
  p2cccd::CandidateDensityStats stats;
  p2cccd::RtCandidateTiming timing;
  if (auto st = p2cccd::ComputeCandidateDensityStats(scene, buffer, records, timing, "cpu_reference", &stats);
      !st.ok) {
    CopyError(st, err, errlen);
    return -3;
  }
  *n_raw = static_cast<int64_t>(buffer.hits.size());
  *n_cand = static_cast<int64_t>(records.size());
  for (size_t i = 0; i < buffer.hits.size() && static_cast<int64_t>(i) < raw_cap; ++i) {
    const auto& h = buffer.hits[i];
    orc_raw_hit& o = raw[i];
    std::memset(&o, 0, sizeof(o));
    o.query_id = h.query_id;
    o.pair_key = h.pair_key;
    o.proxy_a_index = h.proxy_a_index;
    o.proxy_b_index = h.proxy_b_index;
    o.slab_id = h.slab_id;
    o.rt_hit_count = h.rt_hit_count;
    o.flags = h.flags;
    for (int k = 0; k < 4; ++k) o.motion_bound[k] = h.motion_bound[k];
  }
  for (size_t i = 0; i < records.size() && static_cast<int64_t>(i) < cand_cap; ++i) {
    const auto& r = records[i];
    orc_candidate& o = cand[i];
    std::memset(&o, 0, sizeof(o));
    o.candidate_id = r.candidate_id;
    o.query_id = r.query_id;
    o.slab_id = r.slab_id;
    o.object_a_id = r.object_a_id;
    o.patch_a_id = r.patch_a_id;
    o.object_b_id = r.object_b_id;
    o.patch_b_id = r.patch_b_id;
    o.proxy_type_a = static_cast<uint32_t>(r.proxy_type_a);
    o.proxy_type_b = static_cast<uint32_t>(r.proxy_type_b);
    o.rt_hit_count = r.rt_hit_count;
    o.flags = r.flags;
    for (int k = 0; k < 4; ++k) o.motion_bound[k] = r.motion_bound[k];
  }
  density->proxy_count = stats.proxy_count;
  density->object_count = stats.object_count;
  density->slab_count = stats.slab_count;
  density->cross_object_same_slab_pair_count = stats.cross_object_same_slab_pair_count;
  density->raw_hit_count = stats.raw_hit_count;
  density->compact_candidate_count = stats.compact_candidate_count;
  density->raw_hits_per_proxy = stats.raw_hits_per_proxy;
  density->candidates_per_proxy = stats.candidates_per_proxy;
  density->candidates_per_slab = stats.candidates_per_slab;
  density->aabb_overlap_ratio = stats.aabb_overlap_ratio;
  density->avg_rt_hits_per_candidate = stats.avg_rt_hits_per_candidate;
  return 0;
}

// The scene of the reference's candidate KAT (src/tests/cpp/test_rt_candidate.cpp:
// 60-106): two single-patch objects (ids 10 and 20, a swept AABB and a capsule,
// patch radius 0.2, eps_proxy 0.05), static, two slabs each; built by the
// reference's own BuildProxyScene and exported as arrays (n <= cap).
extern "C" int ref_proxy_kat_scene(int64_t cap, uint32_t* object_id, uint32_t* patch_id, uint32_t* slab_id,
                                   uint32_t* proxy_type, double* t0, double* t1, double* bounds, double* motion,
                                   int64_t* n, uint64_t* query_id) {
  auto make_patch = [](uint32_t id, double x) {
    p2cccd::Patch patch;
    patch.patch_id = id;
    patch.triangle_ids = {id};
    patch.triangle_count = 1;
    patch.area = 1.0;
    patch.local_center = {x, 0.0, 0.0};
    patch.radius = 0.2;
    return patch;
  };
  p2cccd::MotionSegment still;
  still.t0 = 0.0;
  still.t1 = 1.0;
  p2cccd::ProxySceneBuildInput input;
  input.query_id = 42;
  p2cccd::ProxyObjectBuildInput a, b;
  a.object_id = 10;
  a.proxy_type = p2cccd::ProxyType::kSweptAabb;
  a.patches = {make_patch(1, 0.0)};
  a.motion_segments = {still};
  a.slabs_per_motion_segment = 2;
  a.eps_proxy = 0.05;
  b = a;
  b.object_id = 20;
  b.proxy_type = p2cccd::ProxyType::kCapsule;
  b.patches = {make_patch(2, 0.25)};
  input.objects = {a, b};
  p2cccd::ProxyScene scene;
  if (!p2cccd::BuildProxyScene(input, &scene).ok) return -1;
  *n = static_cast<int64_t>(scene.primitives.size());
  *query_id = scene.query_id;
  for (int64_t i = 0; i < *n && i < cap; ++i) {
    const auto& p = scene.primitives[static_cast<size_t>(i)];
    object_id[i] = p.object_id;
    patch_id[i] = p.patch_id;
    slab_id[i] = p.slab_id;
    proxy_type[i] = static_cast<uint32_t>(p.proxy_type);
    t0[i] = p.t0;
    t1[i] = p.t1;
    for (int k = 0; k < 3; ++k) {
      bounds[6 * i + k] = p.bounds.min[k];
      bounds[6 * i + 3 + k] = p.bounds.max[k];
    }
    motion[4 * i + 0] = p.motion_bound.translation_bound;
    motion[4 * i + 1] = p.motion_bound.rotation_angle;
    motion[4 * i + 2] = p.motion_bound.radial_motion_bound;
    motion[4 * i + 3] = p.motion_bound.conservative_radius;
  }
  return 0;
}
