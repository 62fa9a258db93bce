// Proxy-scene candidate interface of the reference, kept field for field so
// existing callers drop in, implemented over the gfx950 C-ABI
// (include/p2c_hip.h).  Layouts follow
//   src/cpp/geometry/proxy.h:12-29, src/cpp/geometry/motion_utils.h:19-31,
//   src/cpp/rt_candidate/proxy_scene.h:29-46,
//   src/cpp/rt_candidate/candidate_buffer.h:19-40,
//   src/cpp/rt_candidate/candidate_generation_result.h:12-45,
//   src/cpp/common/runtime_contracts.h:41-56,
//   src/cpp/geometry/motion.h:8-27, src/cpp/geometry/patch.h:8-17,
//   src/cpp/rt_candidate/proxy_scene.h:14-27 (build inputs),
//   src/cpp/rt_candidate/candidate_generator.h:13-21,
//   src/cpp/rt_candidate/external_batch_candidate.h:14-50.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "certificate_hip.h"

namespace p2cccd {

struct Aabb {
  std::array<double, 3> min{0.0, 0.0, 0.0};
  std::array<double, 3> max{0.0, 0.0, 0.0};
};

struct PatchMotionBound {
  std::uint32_t patch_id = 0;
  double t0 = 0.0;
  double t1 = 1.0;
  std::array<double, 3> center_t0{0.0, 0.0, 0.0};
  std::array<double, 3> center_t1{0.0, 0.0, 0.0};
  double translation_bound = 0.0;
  double rotation_angle = 0.0;
  double center_rotation_bound = 0.0;
  double surface_rotation_bound = 0.0;
  double radial_motion_bound = 0.0;
  double conservative_radius = 0.0;
};

struct PoseSample {
  std::array<double, 3> translation{0.0, 0.0, 0.0};
  std::array<double, 4> rotation_xyzw{0.0, 0.0, 0.0, 1.0};
};

struct MotionSegment {
  double t0 = 0.0;
  double t1 = 1.0;
  PoseSample pose_t0;
  PoseSample pose_t1;
};

struct Patch {
  std::uint32_t patch_id = 0;
  std::vector<std::uint32_t> triangle_ids;
  std::uint32_t triangle_count = 0;
  double area = 0.0;
  std::array<double, 3> local_center{0.0, 0.0, 0.0};
  double radius = 0.0;
};

struct TimeSlab {
  std::uint32_t slab_id = 0;
  double t0 = 0.0;
  double t1 = 1.0;
};

struct EndpointSweptAabbProxy {
  std::uint32_t patch_id = 0;
  Aabb bounds;
  PatchMotionBound motion_bound;
};

struct CapsuleProxy {
  std::uint32_t patch_id = 0;
  std::array<double, 3> endpoint0{0.0, 0.0, 0.0};
  std::array<double, 3> endpoint1{0.0, 0.0, 0.0};
  double radius = 0.0;
  PatchMotionBound motion_bound;
};

struct ProxyPrimitive {
  std::uint32_t proxy_id = 0;
  std::uint32_t object_id = 0;
  std::uint32_t patch_id = 0;
  std::uint32_t slab_id = 0;
  std::uint32_t motion_segment_id = 0;
  ProxyType proxy_type = ProxyType::kUnknown;
  double t0 = 0.0;
  double t1 = 1.0;
  Aabb bounds;
  CapsuleProxy capsule;
  PatchMotionBound motion_bound;
};

struct ProxyScene {
  std::uint64_t query_id = 0;
  std::vector<ProxyPrimitive> primitives;
};

struct ProxyObjectBuildInput {
  std::uint32_t object_id = 0;
  ProxyType proxy_type = ProxyType::kSweptAabb;
  std::vector<Patch> patches;
  std::vector<MotionSegment> motion_segments;
  std::uint32_t slabs_per_motion_segment = 1;
  double eps_proxy = 0.0;
};

struct ProxySceneBuildInput {
  std::uint64_t query_id = 0;
  std::vector<ProxyObjectBuildInput> objects;
};

enum RawCandidateFlags : std::uint32_t {
  kRawCandidateValid = 1U << 0U,
  kRawCandidateAabbOverlap = 1U << 1U,
};

struct alignas(16) RawCandidateHit {
  std::uint64_t query_id = 0;
  std::uint64_t pair_key = 0;
  std::uint32_t proxy_a_index = 0;
  std::uint32_t proxy_b_index = 0;
  std::uint32_t slab_id = 0;
  std::uint32_t rt_hit_count = 0;
  std::uint32_t flags = 0;
  std::uint32_t reserved0 = 0;
  std::array<float, 4> motion_bound{0.0f, 0.0f, 0.0f, 0.0f};
  std::uint32_t reserved1 = 0;
  std::uint32_t reserved2 = 0;
};

struct RawCandidateBuffer {
  std::vector<RawCandidateHit> hits;
};

struct CandidateRecord {
  std::uint32_t schema_version = 1;
  std::uint64_t candidate_id = 0;
  std::uint64_t query_id = 0;
  std::uint32_t slab_id = 0;
  std::uint32_t object_a_id = 0;
  std::uint32_t patch_a_id = 0;
  std::uint32_t object_b_id = 0;
  std::uint32_t patch_b_id = 0;
  ProxyType proxy_type_a = ProxyType::kUnknown;
  ProxyType proxy_type_b = ProxyType::kUnknown;
  std::uint32_t rt_hit_count = 0;
  std::array<float, 4> motion_bound{0.0f, 0.0f, 0.0f, 0.0f};
  std::uint32_t proxy_features_offset = 0;
  std::uint32_t flags = 0;
};

struct RtCandidateTiming {
  double build_ms = 0.0;
  double update_ms = 0.0;
  double trace_ms = 0.0;
  double compact_ms = 0.0;
  double stats_ms = 0.0;
  double total_ms = 0.0;
};

struct CandidateDensityStats {
  std::uint32_t schema_version = 1;
  std::uint64_t query_id = 0;
  std::uint64_t proxy_count = 0;
  std::uint64_t object_count = 0;
  std::uint64_t slab_count = 0;
  std::uint64_t cross_object_same_slab_pair_count = 0;
  std::uint64_t raw_hit_count = 0;
  std::uint64_t compact_candidate_count = 0;
  double raw_hits_per_proxy = 0.0;
  double candidates_per_proxy = 0.0;
  double candidates_per_slab = 0.0;
  double aabb_overlap_ratio = 0.0;
  double avg_rt_hits_per_candidate = 0.0;
  RtCandidateTiming timing;
  std::string backend_name;
};

struct CandidateGenerationResult {
  std::string backend_name;
  RawCandidateBuffer raw_buffer;
  std::vector<CandidateRecord> candidates;
  RtCandidateTiming timing;
  CandidateDensityStats density;
};

enum class CandidateBackend { kCpuReference, kOptix };

struct CandidateGeneratorConfig {
  CandidateBackend backend = CandidateBackend::kCpuReference;
  bool allow_optix_cpu_fallback = false;
};

enum class ExternalQueryFamily : std::uint8_t { kVertexFace = 0, kEdgeEdge = 1 };

struct ExternalBatchQuery {
  std::uint64_t source_query_id = 0;
  std::uint32_t source_query_index = 0;
  ExternalQueryFamily family = ExternalQueryFamily::kVertexFace;
  std::array<std::array<double, 3>, 4> vertices_t0{};
  std::array<std::array<double, 3>, 4> vertices_t1{};
  bool has_ground_truth = false;
  bool ground_truth_collides = false;
  bool has_box_pair = false;
  std::array<std::uint32_t, 2> box_pair{1, 2};
};

struct RuntimeQueryIdMapping {
  std::uint64_t source_query_id = 0;
  std::uint64_t runtime_query_id = 0;
};

struct ExternalBatchCandidateResult {
  std::string backend_name;
  RtCandidateTiming timing;
  std::uint64_t primitive_count = 0;
  std::uint64_t raw_hit_count = 0;
  std::uint64_t compact_candidate_count = 0;
  double candidate_recall = 0.0;
  std::vector<CandidateRecord> candidates;
  std::vector<RuntimeQueryIdMapping> runtime_query_ids;
};

// Backend name reported by this implementation.
inline constexpr const char* kHipCandidateBackend = "hip_exact";

// ValidateProxyScene (src/cpp/rt_candidate/proxy_scene.cpp:134-154).
Status ValidateProxyScene(const ProxyScene& scene);
// CandidateGenerator::GenerateCandidates (candidate_generator.cpp:13-33) with
// every backend name running the gfx950 path: raw hits and compaction on the
// device with the CPU reference's FP64 semantics.
Status GenerateCandidatesHip(const ProxyScene& scene, std::uint64_t query_id, CandidateGenerationResult* result);
// CompactRawCandidates (candidate_buffer.cpp:183-228) on the device.
Status CompactRawCandidatesHip(const ProxyScene& scene, const RawCandidateBuffer& raw_buffer,
                               std::vector<CandidateRecord>* candidates);
// ComputeCandidateDensityStats (candidate_stats.cpp:128-170).
Status ComputeCandidateDensityStats(const ProxyScene& scene, const RawCandidateBuffer& raw_buffer,
                                    const std::vector<CandidateRecord>& candidates, const RtCandidateTiming& timing,
                                    const std::string& backend_name, CandidateDensityStats* stats);

// Rigid-patch proxies (src/cpp/geometry/motion_utils.cpp:196-325,
// src/cpp/geometry/proxy.cpp:36-133) and BuildProxyScene
// (src/cpp/rt_candidate/proxy_scene.cpp:156-213), on the host with the
// reference's FP64 operation order, so bounds and motion bounds match bit for
// bit.
Status ValidateMotionSegment(const MotionSegment& segment);
Status GenerateUniformTimeSlabs(double t0, double t1, std::uint32_t slab_count, std::vector<TimeSlab>* slabs);
Status InterpolateRigidMotion(const MotionSegment& segment, double t, PoseSample* pose);
Status TransformPoint(const PoseSample& pose, const std::array<double, 3>& local_point,
                      std::array<double, 3>* world_point);
Status ComputeRotationAngle(const PoseSample& pose0, const PoseSample& pose1, double* angle);
Status ComputePatchMotionBound(const Patch& patch, const MotionSegment& segment, const TimeSlab& slab,
                               PatchMotionBound* bound);
Status BuildEndpointSweptAabbProxy(const Patch& patch, const MotionSegment& segment, const TimeSlab& slab,
                                   double eps_proxy, EndpointSweptAabbProxy* proxy);
Status BuildCapsuleProxy(const Patch& patch, const MotionSegment& segment, const TimeSlab& slab, double eps_proxy,
                         CapsuleProxy* proxy);
Status BuildProxyScene(const ProxySceneBuildInput& input, ProxyScene* scene);

// GenerateCandidatesForExternalBatch (external_batch_candidate.cpp:231-339):
// validation, runtime query ids and the slabs' x shifts on the host in the
// reference's order, then boxes, overlap, candidates and ids on the device
// (p2c_external_batch_host).  Every backend name runs the device path.
Status GenerateCandidatesForExternalBatch(const std::vector<ExternalBatchQuery>& queries,
                                          const CandidateGeneratorConfig& config,
                                          ExternalBatchCandidateResult* result);

}  // namespace p2cccd
