/*
 * p2c_hip.h -- C-ABI of the MI355X (gfx950) proposal-to-certificate library
 * (libp2c_hip.so).  Plain pointers and sizes only; every entry point returns
 * 0 on success or a negative P2C_E* code, with a message available from
 * p2c_last_error() (thread-local).  Device-pointer entry points take
 * caller-owned device buffers and the caller's hipStream_t (passed as void*),
 * enqueue work and return without synchronising unless stated otherwise.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   src/cpp/certificate/certificate_cuda.h:10-26   (batch certificate, cross-check)
 *   src/cpp/certificate/certificate_engine.h:101-127 (interval certificate, queue)
 *   src/cpp/rt_candidate/optix_candidate_tracer.cpp:834-889 (device candidate emit)
 *   src/tools/scene_object_envelope_native_ti_walltime.cpp:256-302, 603-696 (VF/EE sets)
 *   src/tools/tight_inclusion_full_query_benchmark.cpp:375-397 (ticcd call site)
 */
#ifndef P2C_HIP_H
#define P2C_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------ */
#define P2C_OK 0
#define P2C_EINVAL -1      /* invalid argument / config (reference ValueError text) */
#define P2C_EHIP -2        /* HIP runtime failure */
#define P2C_ECAPACITY -3   /* a caller-sized output buffer was too small */
#define P2C_EUNSUPPORTED -4

/* ---- status / family codes (runtime_contracts.h:22-33, certificate_engine.h:12-31) */
#define P2C_STATUS_COLLISION 0
#define P2C_STATUS_SEPARATION 1
#define P2C_STATUS_UNDECIDED 2
#define P2C_STATUS_INVALID 255 /* primitive failed validation (ids < 0 / non-finite) */
#define P2C_FAMILY_POINT_TRIANGLE 1
#define P2C_FAMILY_EDGE_EDGE 2

/* LinearVertexTrajectory (certificate_engine.h:32-36): 56 bytes. */
typedef struct p2c_traj {
  int64_t feature_id;
  double position_t0[3];
  double position_t1[3];
} p2c_traj;

/* PointTriangleIntervalPrimitive / EdgeEdgeIntervalPrimitive
 * (certificate_engine.h:38-54): 240 bytes, the same AoS layout the reference
 * CUDA batch ships (certificate_cuda.cu:37-53).
 *   PT: id_a = point_id, id_b = triangle_id, v = {point, tri_v0, tri_v1, tri_v2}
 *   EE: id_a = edge_a_id, id_b = edge_b_id, v = {edge_a0, edge_a1, edge_b0, edge_b1} */
typedef struct p2c_prim {
  int64_t id_a;
  int64_t id_b;
  p2c_traj v[4];
} p2c_prim;

/* CertificateEngineConfig (certificate_engine.h:56-60).  `flags` (not in
 * the reference) selects the walk: 0 = the CPU engine's
 * (certificate_engine.cpp:300-386), P2C_CERT_CUDA_TWIN = the CUDA batch
 * twin's (certificate_cuda.cu:286-346), which returns UNDECIDED on the first
 * interval in pre-order that may not be split instead of searching on,
 * P2C_CERT_PYTHON_TWIN = the Python twin's (pure_exact_cpu.py:98-244): EE
 * guard abs(denom) > 1e-14, motion radius 0.5 (t1 - t0) (largest displacement
 * of each side, summed), contact at the first smallest sample; its TOI is
 * that sample, t1 otherwise, and pad[0] of each result holds the deepest
 * node visited. */
#define P2C_CERT_CUDA_TWIN 1u
#define P2C_CERT_PYTHON_TWIN 2u
typedef struct p2c_cert_config {
  double eps_time;
  double eps_space;
  uint32_t max_subdivision_depth;
  uint32_t flags;
} p2c_cert_config;

/* PrimitiveIntervalResult (certificate_engine.h:69-81): 64 bytes. */
typedef struct p2c_prim_result {
  uint8_t status;
  uint8_t witness_family;
  uint16_t reason_code;
  uint32_t covered_feature_mask;
  double interval_t0;
  double interval_t1;
  double toi_upper;
  double safe_margin_lb;
  int64_t witness_id_a;
  int64_t witness_id_b;
  uint8_t next_refinement_mode;
  uint8_t pad[3];
  uint32_t nodes; /* subdivision nodes evaluated (work counter, not in the reference) */
} p2c_prim_result;

/* ---- library ------------------------------------------------------------ */
const char* p2c_last_error(void);
int p2c_version(void);
/* 1 when a gfx950 device is visible and the kernels can launch. */
int p2c_device_available(void);
int p2c_stream_synchronize(void* stream);

/* ---- reference-parity interval certificate ------------------------------
 * Replaces EvaluatePointTriangleBatchCuda / EvaluateEdgeEdgeBatchCuda
 * (certificate_cuda.h:11-20).  With config->flags 0, CertificateEngine CPU
 * semantics (certificate_engine.cpp:300-386): depth-first left-then-right
 * subdivision, first collision in that order wins, min margin over separated
 * leaves.  With P2C_CERT_CUDA_TWIN, the CUDA twin's: the same walk, ended
 * by the first collision or by the first interval that may not be split.
 * `prims` and `out` are device pointers of n elements.  A primitive that
 * fails validation gets status P2C_STATUS_INVALID and reason_code 3.
 * `invalid_count` (device uint32, may be NULL) is incremented per such one. */
int p2c_cert_ref_batch(int family, const p2c_prim* prims, int64_t n, double interval_t0,
                       double interval_t1, const p2c_cert_config* config,
                       p2c_prim_result* out, uint32_t* invalid_count, void* stream);

/* Closest distance of each primitive at its own time t[i] (device arrays of
 * n), the distance functions of certificate_engine.cpp:186-198 evaluated on
 * the device; exposed so FP64 lerp, sqrt and division can be checked
 * bit-for-bit against the host.  No reference export corresponds to it. */
int p2c_cert_ref_distance(int family, const p2c_prim* prims, const double* t, int64_t n,
                          double* out, void* stream);

/* Host-buffer twin of p2c_cert_ref_batch with the reference's ownership model
 * (host vectors in and out, certificate_cuda.cu:456-552): allocates, copies,
 * launches, synchronises.  Fails with the reference's validation message when
 * any primitive is invalid. */
int p2c_cert_ref_batch_host(int family, const p2c_prim* prims, int64_t n, double interval_t0,
                            double interval_t1, const p2c_cert_config* config,
                            p2c_prim_result* out);

/* The same certificate evaluated on the host CPU from the same source as the
 * device kernel (the reference's CPU entry points EvaluatePointTriangleInterval
 * / EvaluateEdgeEdgeInterval, certificate_engine.h:101-110). */
int p2c_cert_ref_eval_cpu(int family, const p2c_prim* prim, double interval_t0,
                          double interval_t1, const p2c_cert_config* config,
                          p2c_prim_result* out);

/* ---- exact work queue ------------------------------------------------------
 * Replaces ProcessExactWorkQueueCpu and GenerateConservativeRefinementWorkItems
 * (certificate_engine.cpp:593-720, certificate_engine.h:83-120): every entry's
 * certificate (CertificateEngine::Evaluate, :729-824), its audit rows and its
 * refinement children, computed on the device.
 *
 * One queue entry: an ExactWorkItem (runtime_contracts.h:67-81), its query's
 * CertificateEngineConfig (certificate_engine.h:56-60) and its primitives,
 * [pt_begin, pt_begin + pt_count) of the point-triangle array and
 * [ee_begin, ee_begin + ee_count) of the edge-edge array (ranges may overlap
 * and be shared between entries).  120 bytes. */
typedef struct p2c_queue_entry {
  uint64_t work_item_id;
  uint64_t parent_candidate_id;
  uint64_t query_id;
  uint32_t slab_id;
  uint32_t patch_a_id;
  uint32_t patch_b_id;
  uint32_t feature_family_mask;
  double interval_t0;
  double interval_t1;
  uint32_t topk_feature_ids_offset;
  uint16_t depth;
  uint8_t source;
  uint8_t pad0;
  float priority_score;
  uint32_t max_subdivision_depth;
  double eps_time;
  double eps_space;
  int64_t pt_begin;
  int64_t pt_count;
  int64_t ee_begin;
  int64_t ee_count;
} p2c_queue_entry;

/* CertificateResult (runtime_contracts.h:83-99).  96 bytes. */
typedef struct p2c_certificate {
  uint64_t work_item_id;
  uint64_t query_id;
  double interval_t0;
  double interval_t1;
  double toi_upper;
  double safe_margin_lb;
  int64_t witness_id_a;
  int64_t witness_id_b;
  double eps_time;
  double eps_space;
  uint32_t covered_feature_mask;
  uint16_t reason_code;
  uint8_t status;
  uint8_t witness_family;
  uint8_t next_refinement_mode;
  uint8_t pad[7];
} p2c_certificate;

/* AuditLogRow (runtime_contracts.h:101-114).  80 bytes. */
typedef struct p2c_audit_row {
  uint64_t event_id;
  uint64_t query_id;
  uint64_t candidate_id;
  uint64_t work_item_id;
  double interval_t0;
  double interval_t1;
  uint64_t timestamp_us;
  double aux_value0;
  double aux_value1;
  uint16_t action;
  uint16_t depth;
  uint8_t stage;
  uint8_t pad[3];
} p2c_audit_row;

/* ExactWorkQueueConfig (certificate_engine.h:83-87); walk_flags selects the
 * walk as p2c_cert_config.flags does (0: the CPU engine's). */
typedef struct p2c_queue_config {
  uint64_t first_event_id;
  uint64_t first_timestamp_us;
  uint32_t emit_dequeue_events;
  uint32_t walk_flags;
} p2c_queue_config;

/* ExactRefinementConfig (certificate_engine.h:95-99). */
typedef struct p2c_refine_config {
  uint64_t first_child_work_item_id;
  double min_interval_width;
  uint32_t max_child_depth;
  uint32_t pad;
} p2c_refine_config;

/* status words of p2c_exact_queue_run */
#define P2C_QUEUE_STATUS_WORDS 6
/* Device workspace of a queue of n entries whose entries certify at most
 * pt_pairs point-triangle and ee_pairs edge-edge (entry, primitive) pairs
 * (the sums of pt_count and ee_count bound them). */
size_t p2c_exact_queue_workspace_bytes(int64_t n, int64_t pt_pairs, int64_t ee_pairs);
/* Certifies the n entries (device arrays; pts of n_pt and ees of n_ee
 * primitives), in one pass of device kernels, in the reference's queue
 * order semantics:
 *   certs[i]  entry i's certificate: the PT primitives of an entry whose mask
 *             has bit 0, then its EE primitives (bit 1), the first collision
 *             winning, an invalid config or a reached invalid primitive
 *             giving the invalid-input certificate (MakeInvalidInputCertificate);
 *   rows      with emit_dequeue_events, rows[2i] the EXACT/DEQUEUED row and
 *             rows[2i + 1] the CERTIFY row of entry i, event ids and
 *             timestamps counting up from the config's first ones; without,
 *             rows[i] the CERTIFY row;
 *   children  (refine not NULL) the refinement children of every entry in
 *             queue order, as queue entries with their parent's config and
 *             primitive ranges: entry i's first child id is
 *             first_child_work_item_id plus the number of children of entries
 *             before it (GenerateConservativeRefinementWorkItems per entry
 *             with a running first id).  At most child_cap are written.
 * status (device u64[P2C_QUEUE_STATUS_WORDS]): [0] UINT64_MAX, or the first
 * failure as entry * 64 + code in the reference's check order (decode with
 * p2c_exact_queue_error); [1] the child count; [2] UINT64_MAX, or the
 * position of an entry that repeats an earlier (query_id, work_item_id);
 * [3] non-zero when the pair bounds were too small; [4], [5] the PT and EE
 * pairs the children written would certify (the next pass's bounds).
 * Asynchronous on `stream`. */
int p2c_exact_queue_run(const p2c_queue_entry* entries, int64_t n, const p2c_prim* pts, int64_t n_pt,
                        const p2c_prim* ees, int64_t n_ee, int64_t pt_pairs, int64_t ee_pairs,
                        const p2c_queue_config* config, const p2c_refine_config* refine, void* workspace,
                        size_t workspace_bytes, p2c_certificate* certs, p2c_audit_row* rows,
                        p2c_queue_entry* children, int64_t child_cap, uint64_t* status, void* stream);
/* Turns status words read back from p2c_exact_queue_run into the call's
 * result: P2C_OK, or the error code with the reference's message for the
 * first failure (ValidateExactWorkItem, ValidateAuditLogRow,
 * ValidateCertificateResult, ValidateExactWorkQueueCoverage) in
 * p2c_last_error(). */
int p2c_exact_queue_error(const uint64_t* status_host);
/* Host-buffer form (the reference's host vectors): allocates, copies in,
 * runs, synchronises and copies out.  certs holds n, rows n or 2n, children
 * child_cap entries; *n_children gets the child count (may be NULL). */
int p2c_exact_queue_host(const p2c_queue_entry* entries, int64_t n, const p2c_prim* pts, int64_t n_pt,
                         const p2c_prim* ees, int64_t n_ee, const p2c_queue_config* config,
                         const p2c_refine_config* refine, p2c_certificate* certs, p2c_audit_row* rows,
                         p2c_queue_entry* children, int64_t child_cap, int64_t* n_children);

/* ---- (t,u,v) inclusion certificate --------------------------------------
 * Replaces the Tight-Inclusion call sites ticcd::vertexFaceCCD / edgeEdgeCCD
 * (tight_inclusion_full_query_benchmark.cpp:375-397, with ms, tolerance,
 * t_max = 1, max_itr of :24-40).  The certificate's definition is the
 * library's (FP64 corner enclosure of the multilinear F(t,u,v), forward error
 * bound, toi = least t_lo over done boxes); see the design notes.  Pairs are
 * first expanded into an opaque per-pair workspace of P2C_TUV_PAIR_BYTES
 * each, stored in tiles of 64 pairs: a buffer for n pairs takes
 * p2c_tuv_pairs_bytes(n) = ceil(n / 64) * 64 * P2C_TUV_PAIR_BYTES bytes (and
 * a capacity passed to a setup call counts whole tiles); certification
 * writes one p2c_tuv_result per pair.  A pair whose
 * search exceeds max_itr boxes is appended to the replay list and must be
 * finished with p2c_tuv_replay (the fallback-replay queue). */
#define P2C_TUV_PAIR_BYTES 288
#define P2C_TUV_FLAG_OVER_BUDGET 1u /* search stopped at max_itr; result provisional until replayed */
#define P2C_TUV_FLAG_REPLAYED 2u    /* result produced by the canonical replay */
#define P2C_TUV_FLAG_CAPPED 4u      /* the work cap bound: conservative hit at the cap's t */

typedef struct p2c_tuv_result {
  double toi;        /* earliest certified time; +inf when no hit */
  uint64_t boxes;    /* boxes evaluated */
  uint32_t hit;      /* 1 collision (or conservative cap), 0 separation */
  uint32_t flags;    /* P2C_TUV_FLAG_* */
  uint64_t reserved;
} p2c_tuv_result;

size_t p2c_tuv_pair_bytes(void);
size_t p2c_tuv_pairs_bytes(int64_t n);
size_t p2c_tuv_heap_bytes(int64_t heap_cap);

/* v: device (n, 8, 3) doubles, rows 0-3 the vertices at t=0 (VF: p, f0, f1,
 * f2; EE: a0, a1, b0, b1), rows 4-7 the same at t=1. */
int p2c_tuv_setup_positions(int is_ee, const double* v, int64_t n, double ms, double tolerance, void* pairs,
                            void* stream);
/* Gathered from mesh vertex arrays x0, x1 (device (V, 3) doubles) through
 * idx4 (device (n, 4) int32 vertex ids in the same order as above). */
int p2c_tuv_setup_indexed(int is_ee, const double* x0, const double* x1, const int32_t* idx4, int64_t n,
                          double ms, double tolerance, void* pairs, void* stream);
/* Straight from broad-phase pair lists: VF pairs (vertex, face) with topo =
 * faces (F, 3), or EE pairs (edge i, edge j) with topo = edges (E, 2).  With
 * `totals` (the device array written by p2c_bp_count) the pair count is
 * min(n, totals[is_ee]) and EE records are placed after the VF ones
 * (offset totals[0]); without it, n pairs go to slot out_offset onwards.
 * Records at or beyond slot out_cap are not written. */
int p2c_tuv_setup_pairs(int is_ee, const double* x0, const double* x1, const int32_t* topo, const int32_t* pairs,
                        int64_t n, const uint64_t* totals, int64_t out_offset, int64_t out_cap, double ms,
                        double tolerance, void* tuv_pairs, void* stream);
/* One lane per pair (the plain one-launch path).  With `totals` (device) the
 * pair count is min(n, totals[0] + totals[1]). */
int p2c_tuv_plain(const void* pairs, int64_t n, const uint64_t* totals, double tolerance, int64_t max_itr,
                  p2c_tuv_result* out, uint32_t* replay_list, uint32_t* replay_count, uint32_t replay_cap,
                  void* stream);
/* Batched form, two kernels on `stream`: a persistent lane-refill pass over
 * `grid_blocks` blocks of 256 in which a pair may take up to `light_budget`
 * boxes, then a wave-cooperative pass in which the 64 lanes of a wave share
 * each remaining pair's subdivision tree.  `counters` is 32 zero-initialised
 * device bytes (next pair; replay count at byte 8; heavy count at byte 12;
 * heavy cursor at byte 16).  heavy_list and replay_list hold list_cap
 * entries.  `totals` as above. */
int p2c_tuv_persistent(const void* pairs, int64_t n, const uint64_t* totals, double tolerance, int64_t max_itr,
                       int64_t light_budget, p2c_tuv_result* out, void* counters, uint32_t* heavy_list,
                       uint32_t* replay_list, uint32_t list_cap, int grid_blocks, void* stream);
/* Batched form by rounds: round r walks each still-open pair for at most
 * budgets[r] steps, one lane per pair, saving unfinished walks in `states`
 * (n * p2c_tuv_state_bytes() device bytes) and compacting them into
 * list_a / list_b (list_cap entries each); the last round's survivors go to
 * the wave-cooperative pass.  `counters`: 64 zero-initialised device bytes
 * (replay count at byte 8, heavy count at byte 12).  `budgets` is a host
 * array of n_rounds (1..8) entries. */
size_t p2c_tuv_state_bytes(void);
int p2c_tuv_rounds(const void* pairs, int64_t n, const uint64_t* totals, double tolerance, int64_t max_itr,
                   const uint32_t* budgets, int n_rounds, p2c_tuv_result* out, void* counters, void* states,
                   uint32_t* list_a, uint32_t* list_b, uint32_t* replay_list, uint32_t list_cap, int grid_blocks,
                   void* stream);
/* Batched form by streaming: a chip-sized grid, one lane per pair; idle lanes
 * are refilled together once `refill_at` (1..64) lanes of a wave are idle.
 * Pairs needing more than `light_budget` boxes save their walk in `states`
 * (n * p2c_tuv_state_bytes() bytes) and are listed in `heavy_list`, as are
 * the walks of a wave left with `handover_at` (0..64) or fewer busy lanes
 * once every pair has been claimed; p2c_tuv_wave finishes them.  With
 * `totals`, the VF pairs [0, totals[0]) and the EE pairs after them are
 * claimed through two cursors, each wave starting on the range of its share
 * of the grid, so that waves hold one kind of pair.  `counters`: 64
 * zero-initialised device bytes, laid out as for p2c_tuv_persistent, with the
 * EE range's cursor (u64) at byte 24. */
int p2c_tuv_stream(const void* pairs, int64_t n, const uint64_t* totals, double tolerance, int64_t max_itr,
                   int64_t light_budget, int refill_at, int handover_at, p2c_tuv_result* out, void* counters,
                   void* states, uint32_t* heavy_list, uint32_t list_cap, int grid_blocks, void* stream);
/* Wave-cooperative pass over the pairs listed by p2c_tuv_stream (same
 * counters, states and list): one wave per pair, resuming the saved walks.
 * Pairs over max_itr are appended to `replay_list` for p2c_tuv_replay. */
int p2c_tuv_wave(const void* pairs, void* counters, const void* states, const uint32_t* heavy_list,
                 uint32_t list_cap, double tolerance, int64_t max_itr, p2c_tuv_result* out, uint32_t* replay_list,
                 int grid_blocks, void* stream);
/* The wave pass with splitting (split_at > 0; 0: never).  Waves done with
 * the listed pairs wait for pieces of the others: while waves wait, a wave
 * that has taken split_at boxes on its pair or piece hands every other busy
 * lane's walk (its current box and pending siblings) to the waiting waves,
 * through a queue in `tasks` (task_cap * p2c_tuv_task_bytes() device bytes,
 * task_cap >= 20), in the same kernel; pieces may be cut again.  The pair's
 * best is shared through its `states` slot; split pairs (listed in
 * split_list, list_cap entries, list_cap < 2^31) get their results once all
 * pieces are done.  Pairs whose walks need 64-bit keys (a dimension split
 * more than 31 times) are cut once, after split_at boxes, into pieces run by
 * a closing pass.  The last 12 slots of `tasks` hold the queue's counters and
 * a call number kept from call to call (the buffer is reused; its first
 * contents do not matter); a quarter of the rest holds the 64-bit-key walk's
 * pieces.  When the queue is full, waves keep their work.  `counters` as for
 * p2c_tuv_stream, with the 64-bit-key piece count and cursor at bytes 32 and
 * 44, the piece count at byte 52 (written at the end) and the split count at
 * byte 60.  No counterpart in the reference (its certificate runs one query
 * per CPU thread). */
size_t p2c_tuv_task_bytes(void);
int p2c_tuv_wave_split(const void* pairs, void* counters, void* states, const uint32_t* heavy_list,
                       uint32_t list_cap, double tolerance, int64_t max_itr, p2c_tuv_result* out,
                       uint32_t* replay_list, int64_t split_at, void* tasks, uint32_t task_cap, uint32_t* split_list,
                       int grid_blocks, void* stream);
/* Fallback replay of the listed pairs in canonical order; `heaps` holds
 * slots * p2c_tuv_heap_bytes(heap_cap) bytes of device memory. */
int p2c_tuv_replay(const void* pairs, const uint32_t* replay_list, const uint32_t* replay_count,
                   uint32_t replay_cap, double tolerance, int64_t max_itr, void* heaps, int64_t heap_cap,
                   int64_t slots, p2c_tuv_result* out, void* stream);

/* Compact pair records of the broad-phase path: P2C_TUV_LITE_BYTES per pair
 * (the four vertex ids and the split sequence, in the same 64-pair tiles;
 * p2c_tuv_lite_bytes(n) bytes for n pairs).  A kernel loading a pair
 * rebuilds its positions, displacements and error bounds from the vertex
 * arrays named by `verts`, which must be the arrays and the minimum
 * separation the records were set up with; results are those of the
 * 288-byte records.  p2c_tuv_setup_pairs_lite takes the arguments of
 * p2c_tuv_setup_pairs; the _lite certification calls take those of
 * p2c_tuv_plain, p2c_tuv_stream, p2c_tuv_wave_split and p2c_tuv_replay plus
 * `verts`.  No
 * counterpart in the reference (its certificate reads positions from the
 * meshes per query on the CPU). */
#define P2C_TUV_LITE_BYTES 64
typedef struct p2c_tuv_vertices {
  const double* x0; /* device (V, 3) positions at t = 0 */
  const double* x1; /* device (V, 3) positions at t = 1 */
  double ms;        /* minimum separation */
} p2c_tuv_vertices;
size_t p2c_tuv_lite_bytes(int64_t n);
int p2c_tuv_setup_pairs_lite(int is_ee, const double* x0, const double* x1, const int32_t* topo,
                             const int32_t* pairs, int64_t n, const uint64_t* totals, int64_t out_offset,
                             int64_t out_cap, double ms, double tolerance, void* lite_pairs, void* stream);
/* Compact records of a batch of scenes' pairs (the pair list and offsets of
 * p2c_bp_batch_count / p2c_bp_batch_emit), in one launch.  x0, x1 hold every
 * scene's vertices, scene k's from row scenes[k].vertex_base on; the
 * records name rows of x0 / x1, so the certificate entry points above run on
 * them with verts = {x0, x1, ms}.  Record p is pair p of the list: the same
 * record p2c_tuv_setup_pairs_lite writes for that pair of that scene alone.
 * scenes: device array of n_scenes entries; offsets: device, as written by
 * p2c_bp_batch_count; n: the number of pairs (offsets[2 * n_scenes]).  No
 * counterpart in the reference. */
typedef struct p2c_tuv_scene {
  const int32_t* faces; /* device (F, 3) scene-local vertex ids */
  const int32_t* edges; /* device (E, 2) scene-local vertex ids */
  int64_t vertex_base;  /* first row of the scene in x0 / x1 */
} p2c_tuv_scene;
int p2c_tuv_setup_pairs_lite_batch(const double* x0, const double* x1, const p2c_tuv_scene* scenes,
                                   int32_t n_scenes, const int32_t* pairs, const uint64_t* offsets, int64_t n,
                                   double ms, double tolerance, void* lite_pairs, void* stream);
int p2c_tuv_plain_lite(const void* lite_pairs, const p2c_tuv_vertices* verts, int64_t n, const uint64_t* totals,
                       double tolerance, int64_t max_itr, p2c_tuv_result* out, uint32_t* replay_list,
                       uint32_t* replay_count, uint32_t replay_cap, void* stream);
int p2c_tuv_stream_lite(const void* lite_pairs, const p2c_tuv_vertices* verts, int64_t n, const uint64_t* totals,
                        double tolerance, int64_t max_itr, int64_t light_budget, int refill_at, int handover_at,
                        p2c_tuv_result* out, void* counters, void* states, uint32_t* heavy_list, uint32_t list_cap,
                        int grid_blocks, void* stream);
int p2c_tuv_wave_split_lite(const void* lite_pairs, const p2c_tuv_vertices* verts, void* counters, void* states,
                            const uint32_t* heavy_list, uint32_t list_cap, double tolerance, int64_t max_itr,
                            p2c_tuv_result* out, uint32_t* replay_list, int64_t split_at, void* tasks,
                            uint32_t task_cap, uint32_t* split_list, int grid_blocks, void* stream);
int p2c_tuv_replay_lite(const void* lite_pairs, const p2c_tuv_vertices* verts, const uint32_t* replay_list,
                        const uint32_t* replay_count, uint32_t replay_cap, double tolerance, int64_t max_itr,
                        void* heaps, int64_t heap_cap, int64_t slots, p2c_tuv_result* out, void* stream);

/* Allocation-free certificate.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It certifies n
 * pairs set up in `pairs` (a caller buffer of p2c_tuv_pairs_bytes(n) bytes)
 * exactly as p2c_tuv_stream (light_budget 1024, refill_at 8, handover_at 8,
 * 3 blocks of 256 per CU) + p2c_tuv_wave + p2c_tuv_replay (heap_cap 65,536,
 * 8 slots) do, writing one result per pair to `out`, but keeps every scratch
 * buffer (counters, saved walks, heavy and replay lists, replay heaps) in one
 * static device array of about 30 MB per GPU, held until the process ends;
 * it calls no hipMalloc.  n is at most p2c_tuv_static_max_pairs() (262,144);
 * larger n is refused with P2C_EINVAL.  All calls on one GPU share that
 * array: calls that overlap on one GPU (two streams, two threads) race and
 * must be serialised by the caller. */
size_t p2c_tuv_static_max_pairs(void);
int p2c_tuv_certify_static(const void* pairs, int64_t n, double tolerance, int64_t max_itr, p2c_tuv_result* out,
                           void* stream);

/* ---- swept-AABB LBVH broad phase ----------------------------------------
 * Replaces the OptiX candidate emitter (optix_candidate_tracer.cpp:652-889,
 * optix_candidate_program.cu:66-130) with the candidate-set semantics of the
 * reference's sort-and-sweep (scene_object_envelope_native_ti_walltime.cpp:
 * 256-302 boxes, 528-550 object envelopes, 595-696 VF/EE sets): candidate
 * sets are identical as sets.  All pointers in p2c_bp_mesh are device
 * pointers; topology (faces, edges, objects) is prepared once per mesh as in
 * the reference's PrepareScene (:497-586).  At most 32768 objects per scene
 * (the object-pair matrix takes n_objects^2 bytes of the workspace). */
typedef struct p2c_bp_mesh {
  const int32_t* faces;          /* (F, 3) */
  const int32_t* edges;          /* (E, 2), (min, max) */
  const int32_t* vertex_object;  /* (V,), -1 for vertices in no face */
  const int32_t* face_object;    /* (F,) */
  const int32_t* edge_object;    /* (E,) */
  int64_t n_vertices, n_faces, n_edges;
  int32_t n_objects;
  int32_t reserved;
  /* Optional scratch kept between p2c_bp_count and p2c_bp_emit: int32
   * [(V + E) * hit_cache_slots].  The count pass keeps each query's first
   * hit_cache_slots partners there and the fill copies them, traversing
   * again only for queries with more; NULL (or 0 slots): the fill traverses
   * every query.  Output is the same either way. */
  int32_t* hit_cache;
  int64_t hit_cache_slots;
} p2c_bp_mesh;

size_t p2c_bp_workspace_bytes(int64_t n_vertices, int64_t n_faces, int64_t n_edges, int32_t n_objects);
/* Byte offsets, within a workspace of these sizes, of what the build and
 * count passes leave there (for inspection and measurement): the face and
 * edge trees (2n - 1 nodes of 64 bytes: box min[3], max[3] as doubles, then
 * int32 left, right, escape, prim; internal nodes at [0, n - 1), leaves in
 * tree order at [n - 1, 2n - 1); a walk goes to `left` when a node's box
 * overlaps the query and it has children, else to `escape`, -1 ending it),
 * the vertex and edge query boxes (min[3], max[3] doubles), the vertex and
 * edge query orders (u32; query s is vertex qorder[s] / edge eorder[s]), the
 * per-query candidate counts (u32, vertex queries then edge queries) and
 * the nodes at which the emit pass resumes a query with more candidates
 * than the hit cache holds (int32, same order).  A vertex query walks the
 * face tree from node 0; edge query s walks the edge tree from the escape
 * of leaf s (each unordered pair once).  No counterpart in the reference. */
typedef struct p2c_bp_layout {
  size_t fnodes, enodes, vbox, ebox, qorder, eorder, counts, resume;
} p2c_bp_layout;
int p2c_bp_workspace_layout(int64_t n_vertices, int64_t n_faces, int64_t n_edges, int32_t n_objects,
                            p2c_bp_layout* out);
/* Swept boxes (FP64, +/- thickness), object envelopes, both LBVHs (faces for
 * VF, edges for EE) and the vertex query order.  x0, x1: device (V, 3). */
int p2c_bp_build(const p2c_bp_mesh* mesh, const double* x0, const double* x1, double thickness, void* workspace,
                 size_t workspace_bytes, void* stream);
/* Traversal count pass and scan; writes to the device array totals[3]: VF
 * total, EE total, and the number of traversals that hit the step bound
 * (always 0 for a well-formed tree). */
int p2c_bp_count(const p2c_bp_mesh* mesh, void* workspace, uint64_t* totals, void* stream);
/* Fill pass: VF pairs (vertex, face) and EE pairs (edge i, edge j), i < j,
 * as int32 pairs; entries beyond a capacity are dropped (compare with the
 * totals from p2c_bp_count).  Output order: queries in Morton order, leaves
 * in traversal order. */
int p2c_bp_emit(const p2c_bp_mesh* mesh, void* workspace, int32_t* vf_pairs, int64_t vf_cap, int32_t* ee_pairs,
                int64_t ee_cap, void* stream);

/* ---- batched broad phase over many resident scenes -----------------------
 * No counterpart in the reference (its scene tool runs one scene per
 * process, scene_object_envelope_native_ti_walltime.cpp:1239-1281); it is
 * for serving per-timestep queries of many resident scenes.  Each scene
 * keeps its boxes, envelopes, object-pair matrix and both trees in a
 * workspace of its own (p2c_bp_scene_workspace_bytes); a batch of scenes is
 * built, counted and emitted with one launch per stage for the whole batch
 * (the scenes' blocks side by side in each launch, one segmented sort over
 * every scene's keys), with sort keys, per-query counts and the hit cache in
 * a batch workspace.  For every scene the pairs, and their order, are those
 * p2c_bp_build + p2c_bp_count + p2c_bp_emit give for that scene alone.
 * The hit_cache fields of each mesh are ignored: the batch workspace holds
 * hit_cache_slots partners per query.  At most 1024 scenes per batch, each
 * scene workspace at most once. */
typedef struct p2c_bp_scene {
  p2c_bp_mesh mesh;       /* topology and sizes (device pointers) */
  const double* x0;       /* device (V, 3) positions at t = 0 */
  const double* x1;       /* device (V, 3) positions at t = 1 */
  double thickness;
  void* workspace;        /* device, p2c_bp_scene_workspace_bytes(...) bytes */
  size_t workspace_bytes;
} p2c_bp_scene;

size_t p2c_bp_scene_workspace_bytes(int64_t n_vertices, int64_t n_faces, int64_t n_edges, int32_t n_objects);
size_t p2c_bp_batch_workspace_bytes(const p2c_bp_scene* scenes, int32_t n_scenes, int64_t hit_cache_slots);
/* Build and count every scene of the batch (scenes: host array).  Writes
 * the device array offsets[2 * n_scenes + 2]: scene k's VF pairs go to
 * [offsets[2k], offsets[2k + 1]) of the batch's pair list, its EE pairs to
 * [offsets[2k + 1], offsets[2k + 2]); offsets[2 * n_scenes] is the total and
 * offsets[2 * n_scenes + 1] the number of traversals that hit the step
 * bound (always 0 for well-formed trees). */
int p2c_bp_batch_count(const p2c_bp_scene* scenes, int32_t n_scenes, int64_t hit_cache_slots, void* batch_workspace,
                       size_t batch_workspace_bytes, uint64_t* offsets, void* stream);
/* Fill the batch's pair list (int32 pairs; VF pairs (vertex, face), EE
 * pairs (edge i, edge j), i < j, in scene-local ids, at the offsets of
 * p2c_bp_batch_count); pairs at or beyond `cap` are dropped.  Called after
 * p2c_bp_batch_count with the same scenes, slots and batch workspace. */
int p2c_bp_batch_emit(const p2c_bp_scene* scenes, int32_t n_scenes, int64_t hit_cache_slots, void* batch_workspace,
                      size_t batch_workspace_bytes, int32_t* pairs, int64_t cap, void* stream);

/* ---- proxy-scene raw candidates --------------------------------------------
 * Replaces GenerateRawCandidatesCpu (src/cpp/rt_candidate/candidate_buffer.cpp:
 * 144-181) and the OptiX trace it stands beside, with the CPU path's FP64
 * semantics.  The set holds device arrays of n proxy primitives (proxy_id =
 * index, ProxyPrimitive of src/cpp/rt_candidate/proxy_scene.h:29-41). */
typedef struct {
  const uint32_t* object_id;
  const uint32_t* patch_id;
  const uint32_t* slab_id;
  const uint32_t* proxy_type; /* 1 swept AABB, 2 capsule */
  const double* t0;
  const double* t1;
  const double* bounds; /* [n][6]: min xyz then max xyz */
  const double* motion; /* [n][4]: translation_bound, rotation_angle, radial_motion_bound, conservative_radius */
  int64_t n;
  uint64_t query_id;
} p2c_proxy_set;

/* RawCandidateHit (candidate_buffer.h:19-31), 64 bytes. */
#define P2C_RAW_VALID 1u
#define P2C_RAW_AABB_OVERLAP 2u
typedef struct {
  uint64_t query_id;
  uint64_t pair_key;
  uint32_t proxy_a_index, proxy_b_index, slab_id, rt_hit_count, flags, reserved0;
  float motion_bound[4];
  uint32_t reserved1, reserved2;
} p2c_raw_hit;

/* Device workspace of the raw-hit count and fill: per (column segment, row)
 * counts, 4 bytes x n x segments (at most 64 segments). */
size_t p2c_proxy_raw_workspace_bytes(int64_t n);
/* Raw hits per primitive i (pairs i < j), into counts[n] (device); the
 * workspace keeps the per-segment counts for the fill.  Replaces the pair
 * loop of GenerateRawCandidatesCpu (candidate_buffer.cpp:144-181). */
int p2c_proxy_raw_count(const p2c_proxy_set* set, uint32_t* counts, void* workspace, size_t workspace_bytes,
                        void* stream);
/* Raw hits in the reference's (i, j) order: primitive i's at offsets[i]
 * (device int64, the exclusive scan of the counts), with the workspace the
 * count call filled; hits at or beyond `cap` are dropped. */
int p2c_proxy_raw_fill(const p2c_proxy_set* set, const int64_t* offsets, const void* workspace,
                       size_t workspace_bytes, p2c_raw_hit* out, int64_t cap, void* stream);

/* CandidateRecord (src/cpp/common/runtime_contracts.h:41-56), 80 bytes with
 * the C++ struct's natural alignment (reserved fields are its padding). */
typedef struct {
  uint32_t schema_version, reserved0;
  uint64_t candidate_id;
  uint64_t query_id;
  uint32_t slab_id, object_a_id, patch_a_id, object_b_id, patch_b_id;
  uint8_t proxy_type_a, proxy_type_b;
  uint16_t reserved1;
  uint32_t rt_hit_count;
  float motion_bound[4];
  uint32_t proxy_features_offset;
  uint32_t flags;
  uint32_t reserved2;
} p2c_candidate;

/* Exclusive scan of n u32 values into out[n + 1] (u64; out[n] = total);
 * part: device u64[n / 1024 + 2].  Used for the raw-hit offsets. */
int p2c_exclusive_scan_u32(const uint32_t* in, int64_t n, uint64_t* out, uint64_t* part, void* stream);
/* Workspace bytes of p2c_proxy_compact for n_raw hits. */
size_t p2c_proxy_compact_workspace_bytes(int64_t n_raw);
/* CompactRawCandidates (src/cpp/rt_candidate/candidate_buffer.cpp:183-228)
 * on the device: validates every raw hit as ValidateRawHit (:66-102), merges
 * hits of equal MakeCandidateKey (:104-124) in buffer order (rt_hit_count
 * sum, per-field std::max of the motion bounds, OR of the flags) and writes
 * the candidates in key order with candidate_id 1, 2, ...  Proxy types must
 * be below 256 (the reference's ProxyType is a byte).  status (device
 * u64[2]): [0] the candidate count, [1] 0 on success, else 1 + (index of the
 * first failing hit in buffer order << 4 | code); p2c_proxy_status_message
 * gives the reference's message for it.  Candidates at or beyond `cap` are
 * dropped. */
int p2c_proxy_compact(const p2c_proxy_set* set, const p2c_raw_hit* raw, int64_t n_raw, void* workspace,
                      size_t workspace_bytes, p2c_candidate* out, int64_t cap, uint64_t* status, void* stream);
const char* p2c_proxy_status_message(uint64_t error_word);
/* Host-owned forms for the reference's host-vector API
 * (GenerateCandidatesForProxyScene, py_module.cpp:103-115, and
 * CompactRawCandidates): the set's arrays are host pointers; the library
 * copies them in, runs count, scan, fill and compaction on the device and
 * returns malloc'ed raw hits and candidates (free with p2c_host_free).
 * stage_ms (optional, [2]): device time of the trace and of the compaction. */
int p2c_proxy_generate_host(const p2c_proxy_set* host_set, p2c_raw_hit** raw, int64_t* n_raw, p2c_candidate** cand,
                            int64_t* n_cand, double* stage_ms);
int p2c_proxy_compact_host(const p2c_proxy_set* host_set, const p2c_raw_hit* raw, int64_t n_raw,
                           p2c_candidate** cand, int64_t* n_cand);
void p2c_host_free(void* p);

/* ---- external query batches ------------------------------------------------
 * Replaces GenerateCandidatesForExternalBatch's proxy trace and compaction
 * (src/cpp/rt_candidate/external_batch_candidate.cpp:231-339).  One record
 * per query of the batch (ExternalBatchQuery, external_batch_candidate.h:
 * 20-31), with what the host resolves in the reference's sequential order:
 * the runtime query id, the x shift of the query's slab in the batched proxy
 * scene, and the box pair (1, 2 when the query has none).  224 bytes. */
typedef struct p2c_external_query {
  double vertices_t0[4][3];
  double vertices_t1[4][3];
  double x_offset;
  uint64_t runtime_query_id;
  uint32_t source_query_index;
  uint32_t family; /* 0 vertex-face (points {0} vs {1,2,3}), 1 edge-edge ({0,1} vs {2,3}) */
  uint32_t object_a_id;
  uint32_t object_b_id;
} p2c_external_query;

size_t p2c_external_batch_workspace_bytes(int64_t n);
/* Candidates of n queries on the device: query order[k] (device u32, the
 * queries by ascending runtime id) is the k-th visited; a query whose two
 * shifted swept boxes overlap (inclusive FP64 test) yields one candidate,
 * written in visiting order with candidate_id = runtime_query_id * 1000003 +
 * ordinal + 1 (the reference's MakeCandidateId over its sorted list).  status
 * (device u64[4]): [0] candidate count; [1] the first query (input index)
 * whose shifted boxes are not finite, [2] the first query whose merged motion
 * bound is not a finite float, [3] the first candidate position whose id
 * wraps to 0, each UINT64_MAX when there is none (ValidateProxyScene,
 * ValidateRawHit and ValidateCandidateRecord failures, in that precedence).
 * Candidates at or beyond `cap` are dropped. */
int p2c_external_batch_candidates(const p2c_external_query* queries, const uint32_t* order, int64_t n,
                                  void* workspace, size_t workspace_bytes, p2c_candidate* out, int64_t cap,
                                  uint64_t* status, void* stream);
/* Host-buffer form (the reference's host vectors): copies in, runs, copies
 * status[4] and the candidates back (malloc'ed, free with p2c_host_free);
 * trace_ms (optional) is the device time. */
int p2c_external_batch_host(const p2c_external_query* queries, const uint32_t* order, int64_t n,
                            p2c_candidate** cand, int64_t* n_cand, uint64_t* status, double* trace_ms);

/* ---- scheduling scorers ---------------------------------------------------
 * Scores only order certificate work; exact certificates decide every pair.
 *
 * TinySTPF candidate scores (FP64), replacing ScoreCandidate with its
 * FeatureFromCandidateWithIndex / PredictTinyStpf and GroupKey
 * (scene_object_envelope_native_ti_walltime.cpp:756-771, 893-1131).  Weights
 * are packed as doubles: mean[32] std[32] w0[64][32] b0[64] w1[32][64]
 * b1[32] w2[32] b2 (the order of LoadTinyStpfWeights, :995-1066). */
#define P2C_TINY_STPF_WEIGHTS 4289
#define P2C_SCORE_LEARNED 0
#define P2C_SCORE_LEARNED_RESIDUAL 1
#define P2C_SCORE_PROXIMITY 2
#define P2C_SCORE_MOTION 3
#define P2C_SCORE_RANDOM 4
/* Per candidate c of a VF (family 1: pairs (vertex, face), topo = faces[nf][3],
 * object_a = vertex objects, object_b = face objects) or EE list (family 2:
 * pairs (edge, edge), topo = edges[ne][2], object_a = object_b = edge
 * objects): scores[c] under `policy` with candidate index c, and
 * group_keys[c] = (min object << 32) | max object.  `totals` (device [VF, EE,
 * ...] counts, may be NULL) caps n.  scores or group_keys may be NULL. */
int p2c_tiny_stpf_score(int family, const double* x0, const double* x1, const int32_t* topo, const int32_t* pairs,
                        int64_t n, const uint64_t* totals, const int32_t* object_a, const int32_t* object_b,
                        int policy, const double* weights, double* scores, uint64_t* group_keys, void* stream);
/* The 32 features of n candidates from their rows (double[n][8][3], the 4
 * vertices at t0 then at t1) and candidate indices; out double[n][32]. */
int p2c_tiny_stpf_features(int family, const double* rows, const uint64_t* index, int64_t n, double* out,
                           void* stream);
/* PredictTinyStpf on n feature rows (double[n][32]). */
int p2c_tiny_stpf_predict(const double* weights, const double* features, int64_t n, double* out, void* stream);

/* STPF priority trunk (FP32 weights and accumulation): the trunk and priority
 * head of STPFModel (src/python/p2cccd/proposal/stpf_model.py:45-93) at
 * feature_dim 32, hidden 128, two GELU layers: sigmoid(wp . h2 + bp).  Weights
 * packed as floats: w1[128][32] b1[128] w2[128][128] b2[128] wp[128] bp
 * (torch nn.Linear's [out][in] layout).  features float[n][32], out float[n]. */
#define P2C_STPF_WEIGHTS 20865
int p2c_stpf_priority(const float* weights, const float* features, int64_t n, float* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* P2C_HIP_H */
