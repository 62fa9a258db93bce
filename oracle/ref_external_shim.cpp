// Oracle (test infrastructure only): a C-ABI shim over the reference's own
// GenerateCandidatesForExternalBatch (src/cpp/rt_candidate/
// external_batch_candidate.cpp) and BuildProxyScene (src/cpp/rt_candidate/
// proxy_scene.cpp, with geometry/proxy.cpp and geometry/motion_utils.cpp),
// compiled from the reference sources where they lie into
// oracle/_ref/libp2c_ref.so by oracle/Makefile.  Used in this container to
// write golden vectors (tests/golden/make_external_golden.py).  Never part of
// the shipped library.
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle.h"
#include "rt_candidate/external_batch_candidate.h"
#include "rt_candidate/proxy_scene.h"

namespace {

void CopyError(const p2cccd::Status& st, char* err, int errlen) {
  if (err != nullptr && errlen > 0) {
    std::strncpy(err, st.message.c_str(), static_cast<size_t>(errlen - 1));
    err[errlen - 1] = 0;
  }
}

}  // namespace

extern "C" int ref_external_batch(int64_t n, const orc_ext_query* q, int64_t cap, orc_candidate* cand,
                                  int64_t* n_cand, uint64_t* runtime_ids, orc_ext_stats* stats, char* err,
                                  int errlen) {
  std::vector<p2cccd::ExternalBatchQuery> queries(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; ++i) {
    auto& d = queries[static_cast<size_t>(i)];
    d.source_query_id = q[i].source_query_id;
    d.source_query_index = q[i].source_query_index;
    d.family = q[i].family == 0 ? p2cccd::ExternalQueryFamily::kVertexFace : p2cccd::ExternalQueryFamily::kEdgeEdge;
    for (int v = 0; v < 4; ++v)
      for (int a = 0; a < 3; ++a) {
        d.vertices_t0[v][a] = q[i].v0[v][a];
        d.vertices_t1[v][a] = q[i].v1[v][a];
      }
    d.has_ground_truth = q[i].has_ground_truth != 0;
    d.ground_truth_collides = q[i].ground_truth_collides != 0;
    d.has_box_pair = q[i].has_box_pair != 0;
    d.box_pair = {q[i].box_pair[0], q[i].box_pair[1]};
  }
  p2cccd::CandidateGeneratorConfig config;
  p2cccd::ExternalBatchCandidateResult result;
  if (auto st = p2cccd::GenerateCandidatesForExternalBatch(queries, config, &result); !st.ok) {
    CopyError(st, err, errlen);
    return -1;
  }
  *n_cand = static_cast<int64_t>(result.candidates.size());
  for (int64_t i = 0; i < *n_cand && i < cap; ++i) {
    const auto& c = result.candidates[static_cast<size_t>(i)];
    orc_candidate& o = cand[i];
    o.candidate_id = c.candidate_id;
    o.query_id = c.query_id;
    o.slab_id = c.slab_id;
    o.object_a_id = c.object_a_id;
    o.patch_a_id = c.patch_a_id;
    o.object_b_id = c.object_b_id;
    o.patch_b_id = c.patch_b_id;
    o.proxy_type_a = static_cast<uint32_t>(c.proxy_type_a);
    o.proxy_type_b = static_cast<uint32_t>(c.proxy_type_b);
    o.rt_hit_count = c.rt_hit_count;
    for (int k = 0; k < 4; ++k) o.motion_bound[k] = c.motion_bound[k];
    o.flags = c.flags;
    o.reserved = c.proxy_features_offset;
  }
  for (size_t i = 0; i < result.runtime_query_ids.size() && int64_t(i) < n; ++i)
    runtime_ids[i] = result.runtime_query_ids[i].runtime_query_id;
  stats->primitive_count = result.primitive_count;
  stats->raw_hit_count = result.raw_hit_count;
  stats->compact_candidate_count = result.compact_candidate_count;
  stats->candidate_recall = result.candidate_recall;
  return 0;
}

extern "C" int ref_build_proxy_scene(uint64_t query_id, int n_objects, const orc_build_object* objects,
                                     const orc_build_patch* patches, const orc_build_segment* segments,
                                     int64_t cap, double* prims, int64_t* n_prims, char* err, int errlen) {
  p2cccd::ProxySceneBuildInput input;
  input.query_id = query_id;
  for (int o = 0; o < n_objects; ++o) {
    p2cccd::ProxyObjectBuildInput obj;
    obj.object_id = objects[o].object_id;
    obj.proxy_type = static_cast<p2cccd::ProxyType>(objects[o].proxy_type);
    obj.slabs_per_motion_segment = objects[o].slabs_per_motion_segment;
    obj.eps_proxy = objects[o].eps_proxy;
    for (uint32_t k = 0; k < objects[o].n_patches; ++k) {
      const orc_build_patch& s = patches[objects[o].patch0 + k];
      p2cccd::Patch p;
      p.patch_id = s.patch_id;
      p.local_center = {s.local_center[0], s.local_center[1], s.local_center[2]};
      p.radius = s.radius;
      obj.patches.push_back(p);
    }
    for (uint32_t k = 0; k < objects[o].n_segments; ++k) {
      const orc_build_segment& s = segments[objects[o].segment0 + k];
      p2cccd::MotionSegment m;
      m.t0 = s.t0;
      m.t1 = s.t1;
      m.pose_t0.translation = {s.translation0[0], s.translation0[1], s.translation0[2]};
      m.pose_t1.translation = {s.translation1[0], s.translation1[1], s.translation1[2]};
      m.pose_t0.rotation_xyzw = {s.rotation0[0], s.rotation0[1], s.rotation0[2], s.rotation0[3]};
      m.pose_t1.rotation_xyzw = {s.rotation1[0], s.rotation1[1], s.rotation1[2], s.rotation1[3]};
      obj.motion_segments.push_back(m);
    }
    input.objects.push_back(obj);
  }
  p2cccd::ProxyScene scene;
  if (auto st = p2cccd::BuildProxyScene(input, &scene); !st.ok) {
    CopyError(st, err, errlen);
    return -1;
  }
  *n_prims = static_cast<int64_t>(scene.primitives.size());
  for (int64_t i = 0; i < *n_prims && i < cap; ++i) {
    const auto& p = scene.primitives[static_cast<size_t>(i)];
    double* d = prims + ORC_PRIM_DOUBLES * i;
    const auto& mb = p.motion_bound;
    const double v[ORC_PRIM_DOUBLES] = {
        double(p.proxy_id), double(p.object_id), double(p.patch_id), double(p.slab_id),
        double(p.motion_segment_id), double(static_cast<int>(p.proxy_type)), p.t0, p.t1,
        p.bounds.min[0], p.bounds.min[1], p.bounds.min[2], p.bounds.max[0], p.bounds.max[1], p.bounds.max[2],
        p.capsule.endpoint0[0], p.capsule.endpoint0[1], p.capsule.endpoint0[2],
        p.capsule.endpoint1[0], p.capsule.endpoint1[1], p.capsule.endpoint1[2], p.capsule.radius,
        double(mb.patch_id), mb.t0, mb.t1, mb.center_t0[0], mb.center_t0[1], mb.center_t0[2],
        mb.center_t1[0], mb.center_t1[1], mb.center_t1[2], mb.translation_bound, mb.rotation_angle,
        mb.center_rotation_bound, mb.surface_rotation_bound, mb.radial_motion_bound, mb.conservative_radius};
    std::memcpy(d, v, sizeof(v));
  }
  return 0;
}
