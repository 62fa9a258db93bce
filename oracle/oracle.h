/*
 * CPU oracle for the proposal-to-certificate path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / CPU baseline.
 *
 * Contents (each function cites the reference lines it restates):
 *   ref_cert.c    reference-parity interval certificate (CertificateEngine)
 *   ref_queue.c   exact work queue: aggregation, audit rows, refinement children
 *   tuv_cert.c    (t,u,v) inclusion certificate, canonical best-first order
 *   broadphase.c  swept boxes + sort-and-sweep VF/EE candidate sets + brute force
 *   mlp.c         TinySTPF FP64 scorer and the 32-128-128-1 FP32 group scorer
 *
 * Pinning (how each function is checked against something other than itself):
 *   ref_cert  : golden vectors written by the compiled reference
 *               (oracle/_ref, tests/golden/ref_cert_*.npz) + the reference
 *               KATs (toi 0.25, margin > 1.9).
 *   ref_queue : the compiled reference's ProcessExactWorkQueueCpu and
 *               GenerateConservativeRefinementWorkItems (oracle/_ref) field
 *               for field, and its golden vectors (tests/golden/queue_golden.npz).
 *   tuv_cert  : parity with Tight-Inclusion itself is UNPINNED (library not
 *               vendored); pinned by closed forms, exact-rational brute force
 *               (sympy) on small cases and invariants (zero FN, toi <= exact).
 *   broadphase: brute-force O(n^2) enumeration with the same set rules.
 *   mlp       : torch fp32 reference / closed form.
 */
#ifndef P2C_ORACLE_H
#define P2C_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int64_t feature_id;
  double p0[3];
  double p1[3];
} orc_traj;

typedef struct {
  int64_t id_a;
  int64_t id_b;
  orc_traj v[4];
} orc_prim; /* PT: point, tri v0, v1, v2 ; EE: a0, a1, b0, b1 */

typedef struct {
  int32_t status;
  uint32_t covered_feature_mask;
  double interval_t0;
  double interval_t1;
  double toi_upper;
  double safe_margin_lb;
  int32_t witness_family;
  int32_t reason_code;
  int64_t witness_id_a;
  int64_t witness_id_b;
  int32_t next_refinement_mode;
  int32_t nodes; /* subdivision nodes evaluated (work counter) */
} orc_result;

/* family: 1 = point-triangle, 2 = edge-edge.  Returns 0, or a negative
 * validation code (-1 config, -2 interval, -3 ids, -4 non-finite). */
int orc_ref_eval(int family, const orc_prim* prim, double t0, double t1, double eps_time,
                 double eps_space, int max_depth, orc_result* out);
int orc_ref_eval_batch(int family, const orc_prim* prims, int64_t n, double t0, double t1,
                       double eps_time, double eps_space, int max_depth, orc_result* out);
/* The CUDA batch twin's walk (certificate_cuda.cu:286-346): UNDECIDED on the
 * first interval that may not be split, in place of searching on. */
int orc_ref_eval_twin_batch(int family, const orc_prim* prims, int64_t n, double t0, double t1,
                            double eps_time, double eps_space, int max_depth, orc_result* out);
/* The Python twin (pure_exact_cpu.py:98-308) over [0, 1]: status, toi_upper,
 * safe_margin_lb, nodes (exact evals / 3) and, in reason_code, the deepest
 * depth reached. */
int orc_ref_eval_py_batch(int family, const orc_prim* prims, int64_t n, double eps_time, double eps_space,
                          int max_depth, orc_result* out);
double orc_pt_distance(const orc_prim* prim, double t);

/* Exact work queue (certificate_engine.cpp:593-720).  The entry, certificate
 * and audit-row records have the byte layout of the library's C-ABI records
 * so one test buffer feeds both.  An entry is an ExactWorkItem, its query's
 * config and its primitive ranges into the PT and EE arrays. */
typedef struct {
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
} orc_queue_entry;

typedef struct {
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
} orc_certificate;

typedef struct {
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
} orc_audit_row;

/* ProcessExactWorkQueueCpu restated: 0, or -1 with the reference's message
 * for the first failure in *msg.  rows holds 2n (emit) or n rows. */
int orc_queue_process(const orc_queue_entry* entries, int64_t n, const orc_prim* pts, const orc_prim* ees,
                      uint64_t first_event_id, uint64_t first_timestamp_us, int emit, orc_certificate* certs,
                      orc_audit_row* rows, const char** msg);
/* GenerateConservativeRefinementWorkItems for each entry in queue order, the
 * first child id of each being first_child plus the children before it;
 * children keep their parent's config and ranges.  Returns the child count
 * (children beyond cap are counted, not written). */
int64_t orc_queue_refine(const orc_queue_entry* entries, const orc_certificate* certs, int64_t n,
                         uint64_t first_child, uint32_t max_child_depth, double min_interval_width,
                         orc_queue_entry* out, int64_t cap);
double orc_ee_distance(const orc_prim* prim, double t);

/* (t,u,v) inclusion certificate, canonical sequential order. */
typedef struct {
  int32_t hit;       /* 1 collision (or conservative cap), 0 separation */
  int32_t capped;    /* 1 if the work cap ended the search */
  double toi;        /* earliest certified time (inf when no hit) */
  int64_t boxes;     /* boxes evaluated */
} orc_tuv_result;

int orc_tuv_eval(int is_ee, const double* v8x3 /* [8][3]: 4 at t0 then 4 at t1 */, double ms,
                 double tolerance, double t_max, int64_t max_itr, int64_t heap_cap,
                 orc_tuv_result* out);
int orc_tuv_eval_batch(int is_ee, const double* v, int64_t n, double ms, double tolerance,
                       double t_max, int64_t max_itr, int64_t heap_cap, orc_tuv_result* out,
                       int nthreads);
void orc_tuv_box(int is_ee, const double* v8x3, double tl, double th, double ul, double uh, double vl, double vh,
                 double mn[3], double mx[3], double err[3]);
void orc_tuv_params(int is_ee, const double* v8x3, double tolerance, double err_out[3],
                    double tol_out[3]);

/* broad phase */
void orc_swept_boxes(const double* x0, const double* x1, int64_t nv, const int32_t* faces,
                     int64_t nf, const int32_t* edges, int64_t ne, double thickness,
                     double* vbox, double* fbox, double* ebox);
int64_t orc_sweep_vf(const double* vbox, const int32_t* vobj, int64_t nv, const double* fbox,
                     const int32_t* faces, const int32_t* fobj, int64_t nf,
                     const uint8_t* allowed, int32_t nobj, int32_t* out, int64_t cap);
int64_t orc_sweep_ee(const double* ebox, const int32_t* edges, const int32_t* eobj, int64_t ne,
                     const uint8_t* allowed, int32_t nobj, int32_t* out, int64_t cap);
int64_t orc_brute_vf(const double* vbox, const int32_t* vobj, int64_t nv, const double* fbox,
                     const int32_t* faces, const int32_t* fobj, int64_t nf,
                     const uint8_t* allowed, int32_t nobj, int32_t* out, int64_t cap);
int64_t orc_brute_ee(const double* ebox, const int32_t* edges, const int32_t* eobj, int64_t ne,
                     const uint8_t* allowed, int32_t nobj, int32_t* out, int64_t cap);

/* proxy-scene candidates (proxy.c; reference shim ref_proxy_shim.cpp) */
typedef struct {
  const uint32_t* object_id;
  const uint32_t* patch_id;
  const uint32_t* slab_id;
  const uint32_t* proxy_type; /* 1 swept AABB, 2 capsule */
  const double* t0;
  const double* t1;
  const double* bounds; /* [n][6]: min xyz, max xyz */
  const double* motion; /* [n][4]: translation, rotation angle, radial motion, conservative radius */
  int64_t n;
} orc_proxy_set;

typedef struct { /* RawCandidateHit, 64 bytes */
  uint64_t query_id, pair_key;
  uint32_t proxy_a_index, proxy_b_index, slab_id, rt_hit_count, flags, reserved0;
  float motion_bound[4];
  uint32_t reserved1, reserved2;
} orc_raw_hit;

typedef struct { /* CandidateRecord fields */
  uint64_t candidate_id, query_id;
  uint32_t slab_id, object_a_id, patch_a_id, object_b_id, patch_b_id, proxy_type_a, proxy_type_b, rt_hit_count;
  float motion_bound[4];
  uint32_t flags, reserved;
} orc_candidate;

typedef struct {
  uint64_t proxy_count, object_count, slab_count, cross_object_same_slab_pair_count, raw_hit_count,
      compact_candidate_count;
  double raw_hits_per_proxy, candidates_per_proxy, candidates_per_slab, aabb_overlap_ratio,
      avg_rt_hits_per_candidate;
} orc_density;

int orc_proxy_generate(uint64_t query_id, const orc_proxy_set* set, int64_t raw_cap, orc_raw_hit* raw,
                       int64_t* n_raw, int64_t cand_cap, orc_candidate* cand, int64_t* n_cand,
                       orc_density* density);

/* external query batches and proxy-scene construction (reference shim
   ref_external_shim.cpp; the restatement of the external batch rule is
   oracle.external_batch in __init__.py) */
typedef struct {
  uint64_t source_query_id;
  uint32_t source_query_index, family; /* 0 vf, 1 ee */
  double v0[4][3], v1[4][3];
  uint8_t has_ground_truth, ground_truth_collides, has_box_pair, pad;
  uint32_t box_pair[2];
} orc_ext_query;

typedef struct {
  uint64_t primitive_count, raw_hit_count, compact_candidate_count;
  double candidate_recall;
} orc_ext_stats;

/* one rigid object of a proxy-scene build; patches and segments are indexed
   into shared arrays */
typedef struct {
  uint32_t object_id, proxy_type, slabs_per_motion_segment, n_patches, n_segments, patch0, segment0;
  double eps_proxy;
} orc_build_object;
typedef struct {
  uint32_t patch_id;
  double local_center[3], radius;
} orc_build_patch;
typedef struct {
  double t0, t1, translation0[3], rotation0[4], translation1[3], rotation1[4];
} orc_build_segment;
#define ORC_PRIM_DOUBLES 36

/* scorers (mlp.c) */
uint64_t orc_mix64(uint64_t x);
double orc_unit_hash(uint64_t x);
void orc_tiny_features(int is_vf, const double* rows24, uint64_t index, double f[32]);
double orc_tiny_predict(const double* packed_weights, const double f[32]);
double orc_tiny_score(int policy, int is_vf, const double* rows24, uint64_t index, uint64_t group_key,
                      const double* packed_weights);
void orc_tiny_score_batch(int policy, int is_vf, const double* rows, const uint64_t* index, const uint64_t* keys,
                          int64_t n, const double* packed_weights, double* out);

#ifdef __cplusplus
}
#endif
#endif
