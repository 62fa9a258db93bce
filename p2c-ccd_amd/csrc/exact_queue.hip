// Exact work queue on gfx950: every entry's certificate, its audit rows and
// its refinement children, in device kernels.
//
// Replaces ProcessExactWorkQueueCpu, CertificateEngine::Evaluate and
// GenerateConservativeRefinementWorkItems (src/cpp/certificate/
// certificate_engine.cpp:593-661, 663-720, 729-824).  The reference walks the
// queue in order, one entry and one primitive at a time; here the pass is
//   1. setup, one thread per entry: the entry's checks, its interval and
//      config as a walk config, the number of (entry, primitive) pairs it
//      certifies per family, and its (query_id, work_item_id) key in a hash
//      set (the coverage check's duplicate rule);
//   2. an exclusive scan of the pair counts per family;
//   3. the certificate walk kernels over all pairs of a family at once, each
//      pair with its entry's interval and config;
//   4. aggregation, one thread per entry, in the reference's order: the PT
//      pairs then the EE pairs, the first collision winning; then the
//      certificate's and the rows' checks, the rows at fixed positions (every
//      entry has one or two), and the number of refinement children;
//   5. a scan of the child counts and the children, written in queue order.
// Checks that fail record entry * 64 + code with atomicMin, so the host
// reports the first failure the reference's sequential loop would meet.
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "cert_ref_walk.h"
#include "runtime.h"
#include "sort_scan.h"

namespace p2c {
namespace {

constexpr int kQueueBlock = 256;
constexpr uint32_t kEmptySlot = 0xFFFFFFFFu;
constexpr uint64_t kNone = ~0ull;

// Failure codes, in the order the reference's loop checks them for one entry
// (certificate_engine.cpp:611-657 with validators.cpp:207-315).
enum : uint32_t {
  kErrRange = 1,  // a primitive range outside its array (a C-ABI check, not the reference's)
  kErrWorkItemId,
  kErrParentId,
  kErrQueryId,
  kErrT0Finite,
  kErrT1Finite,
  kErrInterval,
  kErrFamilyMask,
  kErrPriority,
  kErrSource,
  kErrDequeueEvent,
  kErrDequeueTimestamp,
  kErrDepth,  // the walk would go deeper than 63 levels
  kErrToiFinite,
  kErrMarginFinite,
  kErrEpsTimeFinite,
  kErrEpsTimePositive,
  kErrEpsSpaceFinite,
  kErrEpsSpacePositive,
  kErrToiInside,
  kErrWitness,
  kErrMarginNegative,
  kErrCoveredMask,
  kErrReason,
  kErrMode,
  kErrCertifyEvent,
  kErrCertifyTimestamp,
  kErrAux0,
  kErrCodeCount
};

enum : uint8_t { kRefineNone = 0, kRefineBisect = 1, kRefineGeometry = 2, kRefineEscalate = 3 };
enum : uint16_t { kReasonNone = 0, kReasonMissingGeometry = 1, kReasonMaxDepth = 2, kReasonInvalidInput = 3 };
enum : uint16_t { kAuditDequeued = 1, kAuditCollision = 2, kAuditSeparation = 3, kAuditUndecided = 4,
                  kAuditInvalidInput = 5 };
enum : uint8_t { kStageExact = 2, kStageCertify = 4 };

// ValidateExactWorkItem (validators.cpp:207-230): 0 or the failing check.
__device__ uint32_t work_item_code(const p2c_queue_entry& e) {
  if (e.work_item_id == 0) return kErrWorkItemId;
  if (e.parent_candidate_id == 0) return kErrParentId;
  if (e.query_id == 0) return kErrQueryId;
  if (!__builtin_isfinite(e.interval_t0)) return kErrT0Finite;
  if (!__builtin_isfinite(e.interval_t1)) return kErrT1Finite;
  if (e.interval_t0 < 0.0 || e.interval_t1 > 1.0 || e.interval_t0 > e.interval_t1) return kErrInterval;
  if (e.feature_family_mask == 0) return kErrFamilyMask;
  if (!__builtin_isfinite(e.priority_score)) return kErrPriority;
  if (e.source > 2) return kErrSource;
  return 0;
}

// ValidateConfig (certificate_engine.cpp:224-232).
__device__ bool config_ok(const p2c_queue_entry& e) {
  return __builtin_isfinite(e.eps_time) && e.eps_time > 0.0 && __builtin_isfinite(e.eps_space) &&
         e.eps_space > 0.0;
}

// The walk's depth bound for the entry's interval and config, the rule of the
// per-batch entry points: false when a node deeper than 63 levels may be
// reached (then the call is refused).
__device__ bool depth_ok(const p2c_queue_entry& e, uint32_t* walk_depth) {
  const double top = fmax(fabs(e.interval_t0), fabs(e.interval_t1));
  const double spacing = nextafter(top, __builtin_huge_val()) - top;
  uint32_t reach = e.max_subdivision_depth;
  if (e.eps_time >= 4.0 * spacing) {
    uint32_t d = 0;
    double w = e.interval_t1 - e.interval_t0;
    while (w > e.eps_time && d < 80) {
      w *= 0.5;
      ++d;
    }
    if (d + 2 < reach) reach = d + 2;
  }
  *walk_depth = e.max_subdivision_depth < 63u ? e.max_subdivision_depth : 63u;
  return reach <= 63;
}

__device__ bool range_ok(int64_t begin, int64_t count, int64_t n) {
  return begin >= 0 && count >= 0 && count <= int64_t(0xFFFFFFFFu) && begin <= n && count <= n - begin;
}

__device__ uint64_t key_hash(uint64_t q, uint64_t w) {
  uint64_t h = q * 0x9E3779B97F4A7C15ull ^ (w + 0x632BE59BD9B4E019ull + (q << 6) + (q >> 2));
  h ^= h >> 31;
  h *= 0xD6E8FEB86659FD93ull;
  h ^= h >> 32;
  return h;
}

__device__ void fail(uint64_t* status, int64_t entry, uint32_t code) {
  atomicMin(reinterpret_cast<unsigned long long*>(&status[0]), (unsigned long long)(uint64_t(entry) * 64 + code));
}

__global__ __launch_bounds__(kQueueBlock) void queue_init_kernel(uint64_t* status, uint32_t* table, int64_t slots) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i == 0) {
    status[0] = kNone;
    status[1] = 0;
    status[2] = kNone;
    status[3] = 0;
    status[4] = 0;
    status[5] = 0;
  }
  for (int64_t k = i; k < slots; k += int64_t(gridDim.x) * blockDim.x) table[k] = kEmptySlot;
}

// Step 1.  An entry certifies its PT pairs when its checks pass, its config
// is valid, its mask has bit 0 and its walk fits 63 levels; likewise EE.
__global__ __launch_bounds__(kQueueBlock) void queue_setup_kernel(const p2c_queue_entry* __restrict__ entries,
                                                                  int64_t n, int64_t n_pt, int64_t n_ee,
                                                                  RefCfg* __restrict__ cfg,
                                                                  uint32_t* __restrict__ cnt_pt,
                                                                  uint32_t* __restrict__ cnt_ee,
                                                                  uint32_t* __restrict__ table, int64_t slots,
                                                                  uint64_t* __restrict__ status) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const p2c_queue_entry e = entries[i];
    uint32_t np = 0, ne = 0;
    const bool ranges = range_ok(e.pt_begin, e.pt_count, n_pt) && range_ok(e.ee_begin, e.ee_count, n_ee);
    if (!ranges) fail(status, i, kErrRange);
    uint32_t walk_depth = 0;
    if (ranges && work_item_code(e) == 0 && config_ok(e)) {
      const bool pt_on = (e.feature_family_mask & 1u) != 0 && e.pt_count > 0;
      const bool ee_on = (e.feature_family_mask & 2u) != 0 && e.ee_count > 0;
      const bool fits = depth_ok(e, &walk_depth);
      if ((pt_on || ee_on) && !fits) fail(status, i, kErrDepth);
      if (fits) {
        np = pt_on ? uint32_t(e.pt_count) : 0u;
        ne = ee_on ? uint32_t(e.ee_count) : 0u;
      }
    }
    cfg[i] = RefCfg{e.interval_t0, e.interval_t1, e.eps_time, e.eps_space, walk_depth};
    cnt_pt[i] = np;
    cnt_ee[i] = ne;
    // duplicate (query_id, work_item_id): open addressing over entry indices;
    // the entries are read-only, so a claimed slot's key is always readable
    const int64_t mask = slots - 1;
    int64_t s = int64_t(key_hash(e.query_id, e.work_item_id) & uint64_t(mask));
    for (;;) {
      const uint32_t old = atomicCAS(&table[s], kEmptySlot, uint32_t(i));
      if (old == kEmptySlot) break;
      const p2c_queue_entry& o = entries[old];
      if (o.query_id == e.query_id && o.work_item_id == e.work_item_id) {
        const int64_t later = int64_t(old) > i ? int64_t(old) : i;
        atomicMin(reinterpret_cast<unsigned long long*>(&status[2]), (unsigned long long)later);
        break;
      }
      s = (s + 1) & mask;
    }
  }
}

// Step 3's source: slot j of a family's launch is pair j of the scan, i.e.
// primitive begin + (j - off[i]) of the entry i with off[i] <= j < off[i + 1],
// with that entry's interval and config.
struct QueueSrc {
  const p2c_prim* prims;
  const p2c_queue_entry* entries;
  const RefCfg* cfg;
  const uint64_t* off;  // n + 1
  int64_t n;
  int64_t cap;
  bool pt;
  __device__ bool at(int64_t j, const p2c_prim** p, RefCfg* c) const {
    const uint64_t total = off[n];
    if (uint64_t(j) >= total || j >= cap) return false;
    int64_t lo = 0, hi = n;  // off[lo] <= j < off[hi]
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (off[mid] <= uint64_t(j)) lo = mid;
      else hi = mid;
    }
    const int64_t begin = pt ? entries[lo].pt_begin : entries[lo].ee_begin;
    *p = prims + begin + (j - int64_t(off[lo]));
    *c = cfg[lo];
    return true;
  }
};

// MakeBaseCertificateResult (certificate_engine.cpp:388-402).
__device__ void base_certificate(const p2c_queue_entry& e, p2c_certificate* r) {
  r->work_item_id = e.work_item_id;
  r->query_id = e.query_id;
  r->interval_t0 = e.interval_t0;
  r->interval_t1 = e.interval_t1;
  r->toi_upper = e.interval_t1;
  r->safe_margin_lb = 0.0;
  r->witness_id_a = -1;
  r->witness_id_b = -1;
  r->eps_time = e.eps_time;
  r->eps_space = e.eps_space;
  r->covered_feature_mask = 0;
  r->reason_code = kReasonMissingGeometry;
  r->status = P2C_STATUS_UNDECIDED;
  r->witness_family = 0;
  r->next_refinement_mode = kRefineGeometry;
  for (int k = 0; k < 7; ++k) r->pad[k] = 0;
}

// RefinementModeForReason (certificate_engine.cpp:368-378).
__device__ uint8_t mode_for_reason(uint16_t reason) {
  switch (reason) {
    case kReasonMissingGeometry:
      return kRefineGeometry;
    case kReasonMaxDepth:
      return kRefineBisect;
    case kReasonInvalidInput:
      return kRefineGeometry;
    default:
      return kRefineEscalate;
  }
}

// ApplyUndecided (certificate_engine.cpp:417-424).
__device__ void apply_undecided(uint16_t reason, p2c_certificate* r) {
  r->status = P2C_STATUS_UNDECIDED;
  r->covered_feature_mask = 0;
  r->safe_margin_lb = 0.0;
  r->reason_code = reason == kReasonNone ? uint16_t(kReasonMaxDepth) : reason;
  r->next_refinement_mode = mode_for_reason(r->reason_code);
}

// ApplyPrimitiveCollision (certificate_engine.cpp:404-415).
__device__ void apply_collision(const p2c_prim_result& p, p2c_certificate* r) {
  r->status = P2C_STATUS_COLLISION;
  r->toi_upper = p.toi_upper;
  r->safe_margin_lb = 0.0;
  r->witness_family = p.witness_family;
  r->witness_id_a = p.witness_id_a;
  r->witness_id_b = p.witness_id_b;
  r->covered_feature_mask = 0;
  r->reason_code = kReasonNone;
  r->next_refinement_mode = kRefineNone;
}

// ValidateCertificateResult (validators.cpp:232-293) on a certificate the
// aggregation built: ids, status, interval and modes hold by construction.
__device__ uint32_t certificate_code(const p2c_certificate& c) {
  if (!__builtin_isfinite(c.toi_upper)) return kErrToiFinite;
  if (!__builtin_isfinite(c.safe_margin_lb)) return kErrMarginFinite;
  if (!__builtin_isfinite(c.eps_time)) return kErrEpsTimeFinite;
  if (!(c.eps_time > 0.0)) return kErrEpsTimePositive;
  if (!__builtin_isfinite(c.eps_space)) return kErrEpsSpaceFinite;
  if (!(c.eps_space > 0.0)) return kErrEpsSpacePositive;
  if (c.status == P2C_STATUS_COLLISION) {
    if (c.toi_upper < c.interval_t0 || c.toi_upper > c.interval_t1) return kErrToiInside;
    if (c.witness_id_a < 0 || c.witness_id_b < 0) return kErrWitness;
  }
  if (c.status == P2C_STATUS_SEPARATION) {
    if (c.safe_margin_lb < 0.0) return kErrMarginNegative;
    if (c.covered_feature_mask == 0) return kErrCoveredMask;
  }
  if (c.status == P2C_STATUS_UNDECIDED && c.reason_code == 0) return kErrReason;
  if (c.status == P2C_STATUS_UNDECIDED && c.next_refinement_mode == kRefineNone) return kErrMode;
  return 0;
}

// AuditActionForCertificate (certificate_engine.cpp:436-448).
__device__ uint16_t audit_action(const p2c_certificate& c) {
  if (c.status == P2C_STATUS_COLLISION) return kAuditCollision;
  if (c.status == P2C_STATUS_SEPARATION) return kAuditSeparation;
  return c.reason_code == kReasonInvalidInput ? kAuditInvalidInput : kAuditUndecided;
}

// MakeAuditRow (certificate_engine.cpp:450-468).
__device__ void make_row(uint64_t event, uint64_t ts, const p2c_queue_entry& e, uint8_t stage, uint16_t action,
                         double aux0, double aux1, p2c_audit_row* r) {
  r->event_id = event;
  r->query_id = e.query_id;
  r->candidate_id = e.parent_candidate_id;
  r->work_item_id = e.work_item_id;
  r->interval_t0 = e.interval_t0;
  r->interval_t1 = e.interval_t1;
  r->timestamp_us = ts;
  r->aux_value0 = aux0;
  r->aux_value1 = aux1;
  r->action = action;
  r->depth = e.depth;
  r->stage = stage;
  r->pad[0] = r->pad[1] = r->pad[2] = 0;
}

// The children GenerateConservativeRefinementWorkItems (:663-720) makes for
// a parent and its certificate: 0, 1 (the collision prefix [t0, clamp(toi)])
// or 2 (the two halves of an undecided interval asking for bisection).
__device__ int child_intervals(const p2c_queue_entry& e, const p2c_certificate& c, const p2c_refine_config& rc,
                               double t[3]) {
  if (uint32_t(e.depth) >= rc.max_child_depth) return 0;
  if (c.status == P2C_STATUS_COLLISION) {
    t[0] = e.interval_t0;
    t[1] = std_clamp(c.toi_upper, e.interval_t0, e.interval_t1);
    return t[1] - t[0] > rc.min_interval_width ? 1 : 0;
  }
  if (c.status == P2C_STATUS_UNDECIDED && c.next_refinement_mode == kRefineBisect) {
    if (e.interval_t1 - e.interval_t0 > rc.min_interval_width) {
      t[0] = e.interval_t0;
      t[1] = 0.5 * (e.interval_t0 + e.interval_t1);
      t[2] = e.interval_t1;
      return 2;
    }
  }
  return 0;
}

// Step 4.  Entries that failed a check before Evaluate get no certificate
// (the reference stops there); the host reports the first failure.
__global__ __launch_bounds__(kQueueBlock) void queue_aggregate_kernel(
    const p2c_queue_entry* __restrict__ entries, int64_t n, const p2c_prim_result* __restrict__ res_pt,
    const p2c_prim_result* __restrict__ res_ee, const uint32_t* __restrict__ cnt_pt,
    const uint32_t* __restrict__ cnt_ee, const uint64_t* __restrict__ off_pt,
    const uint64_t* __restrict__ off_ee, int64_t pt_cap, int64_t ee_cap, p2c_queue_config qc, int refine,
    p2c_refine_config rc, p2c_certificate* __restrict__ certs, p2c_audit_row* __restrict__ rows,
    uint32_t* __restrict__ child_cnt, uint64_t* __restrict__ status) {
  if (blockIdx.x == 0 && threadIdx.x == 0 && (off_pt[n] > uint64_t(pt_cap) || off_ee[n] > uint64_t(ee_cap)))
    status[3] = 1;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const p2c_queue_entry e = entries[i];
    if (child_cnt) child_cnt[i] = 0;
    if (uint32_t code = work_item_code(e)) {
      fail(status, i, code);
      continue;
    }
    const bool emit = qc.emit_dequeue_events != 0;
    const uint64_t k = emit ? 2 * uint64_t(i) : uint64_t(i);
    uint64_t event = qc.first_event_id + k, ts = qc.first_timestamp_us + k;
    if (emit) {
      // ValidateAuditLogRow (validators.cpp:295-315): query id, stage,
      // interval and the finite aux values hold for a valid work item
      if (event == 0) {
        fail(status, i, kErrDequeueEvent);
        continue;
      }
      if (ts == 0) {
        fail(status, i, kErrDequeueTimestamp);
        continue;
      }
      make_row(event, ts, e, kStageExact, kAuditDequeued, double(e.priority_score), double(e.feature_family_mask),
               &rows[k]);
      ++event;
      ++ts;
    }
    p2c_certificate c;
    base_certificate(e, &c);
    bool invalid = !config_ok(e);
    const uint64_t p0 = off_pt[i], e0 = off_ee[i];
    const bool pt_on = (e.feature_family_mask & 1u) != 0, ee_on = (e.feature_family_mask & 2u) != 0;
    if (!invalid && ((pt_on && int64_t(cnt_pt[i]) != e.pt_count) || (ee_on && int64_t(cnt_ee[i]) != e.ee_count)))
      continue;  // refused at setup (range or depth), failure recorded there
    if ((pt_on && p0 + uint64_t(e.pt_count) > uint64_t(pt_cap)) ||
        (ee_on && e0 + uint64_t(e.ee_count) > uint64_t(ee_cap)))
      continue;  // pair bounds too small (status[3]); no result to read
    double margin = __builtin_inf();
    uint32_t covered = 0;
    bool any = false, undecided = false, done = false;
    uint16_t reason = kReasonNone;
    for (int f = 0; f < 2 && !invalid && !done; ++f) {
      const bool on = f == 0 ? pt_on : ee_on;
      if (!on) continue;
      const int64_t count = f == 0 ? e.pt_count : e.ee_count;
      const p2c_prim_result* r = f == 0 ? res_pt + p0 : res_ee + e0;
      if (count == 0) {
        undecided = true;
        reason = kReasonMissingGeometry;
      }
      for (int64_t j = 0; j < count; ++j) {
        any = true;
        const p2c_prim_result p = r[j];
        if (p.status == P2C_STATUS_INVALID) {  // Evaluate fails: the invalid-input certificate
          invalid = true;
          break;
        }
        if (p.status == P2C_STATUS_COLLISION) {
          apply_collision(p, &c);
          done = true;
          break;
        }
        if (p.status == P2C_STATUS_SEPARATION) {
          covered |= p.covered_feature_mask;
          margin = std_min(margin, p.safe_margin_lb);
        } else {
          undecided = true;
          reason = p.reason_code;
        }
      }
    }
    if (invalid) {
      base_certificate(e, &c);
      apply_undecided(kReasonInvalidInput, &c);
    } else if (!done) {
      if (!any) {
        apply_undecided(kReasonMissingGeometry, &c);
      } else if (undecided) {
        apply_undecided(reason, &c);
      } else {
        c.status = P2C_STATUS_SEPARATION;
        c.safe_margin_lb = __builtin_isfinite(margin) ? margin : 0.0;
        c.covered_feature_mask = covered;
        c.reason_code = kReasonNone;
        c.next_refinement_mode = kRefineNone;
      }
    }
    if (uint32_t code = certificate_code(c)) {
      fail(status, i, code);
      continue;
    }
    const double aux0 = c.status == P2C_STATUS_SEPARATION ? c.safe_margin_lb : c.toi_upper;
    if (event == 0) {
      fail(status, i, kErrCertifyEvent);
      continue;
    }
    if (ts == 0) {
      fail(status, i, kErrCertifyTimestamp);
      continue;
    }
    if (!__builtin_isfinite(aux0)) {
      fail(status, i, kErrAux0);
      continue;
    }
    certs[i] = c;
    make_row(event, ts, e, kStageCertify, audit_action(c), aux0, double(c.next_refinement_mode),
             &rows[emit ? k + 1 : k]);
    if (refine) {
      double t[3];
      child_cnt[i] = uint32_t(child_intervals(e, c, rc, t));
    }
  }
}

// Step 5: entry i's children at child_off[i], ids first + child_off[i] + k.
__global__ __launch_bounds__(kQueueBlock) void queue_children_kernel(
    const p2c_queue_entry* __restrict__ entries, int64_t n, const p2c_certificate* __restrict__ certs,
    const uint32_t* __restrict__ child_cnt, const uint64_t* __restrict__ child_off, p2c_refine_config rc,
    p2c_queue_entry* __restrict__ children, int64_t cap, uint64_t* __restrict__ status) {
  if (blockIdx.x == 0 && threadIdx.x == 0) status[1] = child_off[n];
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    if (child_cnt[i] == 0) continue;
    const p2c_queue_entry e = entries[i];
    double t[3];
    const int k = child_intervals(e, certs[i], rc, t);
    const uint64_t np = (e.feature_family_mask & 1u) ? uint64_t(e.pt_count) : 0,
                   ne = (e.feature_family_mask & 2u) ? uint64_t(e.ee_count) : 0;
    if (np) atomicAdd(reinterpret_cast<unsigned long long*>(&status[4]), (unsigned long long)(np * k));
    if (ne) atomicAdd(reinterpret_cast<unsigned long long*>(&status[5]), (unsigned long long)(ne * k));
    for (int c = 0; c < k; ++c) {
      const uint64_t at = child_off[i] + uint64_t(c);
      if (at >= uint64_t(cap)) break;
      p2c_queue_entry ch = e;
      ch.work_item_id = rc.first_child_work_item_id + at;
      ch.interval_t0 = t[c];
      ch.interval_t1 = t[c + 1];
      ch.depth = uint16_t(e.depth + 1);
      ch.source = 1;  // ProposalSource::kRefined
      children[at] = ch;
    }
  }
}

size_t align_up(size_t v) { return (v + 255) & ~size_t(255); }

int64_t table_slots(int64_t n) {
  int64_t s = 64;
  while (s < 2 * n) s <<= 1;
  return s;
}

struct QueueLayout {
  size_t cfg, cnt_pt, cnt_ee, cnt_child, off_pt, off_ee, off_child, part, table, res_pt, res_ee, total;
};

QueueLayout queue_layout(int64_t n, int64_t pt_pairs, int64_t ee_pairs) {
  QueueLayout l;
  size_t at = 0;
  auto take = [&](size_t bytes) {
    const size_t here = at;
    at = align_up(at + bytes);
    return here;
  };
  const size_t un = size_t(n);
  l.cfg = take(sizeof(RefCfg) * un);
  l.cnt_pt = take(sizeof(uint32_t) * un);
  l.cnt_ee = take(sizeof(uint32_t) * un);
  l.cnt_child = take(sizeof(uint32_t) * un);
  l.off_pt = take(sizeof(uint64_t) * (un + 1));
  l.off_ee = take(sizeof(uint64_t) * (un + 1));
  l.off_child = take(sizeof(uint64_t) * (un + 1));
  l.part = take(sizeof(uint64_t) * (un / 1024 + 2));
  l.table = take(sizeof(uint32_t) * size_t(table_slots(n)));
  l.res_pt = take(sizeof(p2c_prim_result) * size_t(pt_pairs > 0 ? pt_pairs : 0));
  l.res_ee = take(sizeof(p2c_prim_result) * size_t(ee_pairs > 0 ? ee_pairs : 0));
  l.total = at;
  return l;
}

const char* queue_message(uint32_t code) {
  switch (code) {
    case kErrRange: return "exact work queue entry primitive range lies outside its primitive array";
    case kErrWorkItemId: return "ExactWorkItem.work_item_id is required";
    case kErrParentId: return "ExactWorkItem.parent_candidate_id is required";
    case kErrQueryId: return "ExactWorkItem.query_id is required";
    case kErrT0Finite: return "interval_t0 must be finite";
    case kErrT1Finite: return "interval_t1 must be finite";
    case kErrInterval: return "interval must satisfy 0 <= interval_t0 <= interval_t1 <= 1";
    case kErrFamilyMask: return "ExactWorkItem.feature_family_mask is required";
    case kErrPriority: return "ExactWorkItem.priority_score must be finite";
    case kErrSource: return "ExactWorkItem.source is invalid";
    case kErrDequeueEvent:
    case kErrCertifyEvent: return "AuditLogRow.event_id is required";
    case kErrDequeueTimestamp:
    case kErrCertifyTimestamp: return "AuditLogRow.timestamp_us is required";
    case kErrDepth: return "subdivision deeper than 63 levels is not supported by the HIP certificate";
    case kErrToiFinite: return "CertificateResult.toi_upper must be finite";
    case kErrMarginFinite: return "CertificateResult.safe_margin_lb must be finite";
    case kErrEpsTimeFinite: return "CertificateResult.eps_time must be finite";
    case kErrEpsTimePositive: return "CertificateResult.eps_time must be positive";
    case kErrEpsSpaceFinite: return "CertificateResult.eps_space must be finite";
    case kErrEpsSpacePositive: return "CertificateResult.eps_space must be positive";
    case kErrToiInside: return "CertificateResult.toi_upper must lie inside the certified interval";
    case kErrWitness: return "CertificateResult collision witnesses are required";
    case kErrMarginNegative: return "CertificateResult.safe_margin_lb must be non-negative";
    case kErrCoveredMask: return "CertificateResult.covered_feature_mask is required for separation";
    case kErrReason: return "CertificateResult.reason_code is required for undecided results";
    case kErrMode: return "CertificateResult.next_refinement_mode is required for undecided results";
    case kErrAux0: return "AuditLogRow.aux_value0 must be finite";
    default: return "exact work queue failed";
  }
}

int check_queue_args(const p2c_queue_config* qc, const p2c_refine_config* rc) {
  if (qc == nullptr) return set_error(P2C_EINVAL, "ExactWorkQueueConfig is required");
  // ValidateQueueConfig (certificate_engine.cpp:426-434)
  if (qc->first_event_id == 0) return set_error(P2C_EINVAL, "ExactWorkQueueConfig.first_event_id must be non-zero");
  if (qc->first_timestamp_us == 0)
    return set_error(P2C_EINVAL, "ExactWorkQueueConfig.first_timestamp_us must be non-zero");
  if (qc->walk_flags & ~(P2C_CERT_CUDA_TWIN | P2C_CERT_PYTHON_TWIN))
    return set_error(P2C_EINVAL, "unknown certificate flags");
  if ((qc->walk_flags & P2C_CERT_CUDA_TWIN) && (qc->walk_flags & P2C_CERT_PYTHON_TWIN))
    return set_error(P2C_EINVAL, "choose one certificate walk: CUDA twin or Python twin");
  if (rc != nullptr) {
    // GenerateConservativeRefinementWorkItems' config checks (:682-688)
    if (rc->first_child_work_item_id == 0)
      return set_error(P2C_EINVAL, "ExactRefinementConfig.first_child_work_item_id must be non-zero");
    if (!std::isfinite(rc->min_interval_width) || rc->min_interval_width <= 0.0)
      return set_error(P2C_EINVAL, "ExactRefinementConfig.min_interval_width must be finite and positive");
  }
  return P2C_OK;
}

}  // namespace
}  // namespace p2c

extern "C" {

size_t p2c_exact_queue_workspace_bytes(int64_t n, int64_t pt_pairs, int64_t ee_pairs) {
  if (n < 0) n = 0;
  return p2c::queue_layout(n, pt_pairs, ee_pairs).total;
}

int p2c_exact_queue_run(const p2c_queue_entry* entries, int64_t n, const p2c_prim* pts, int64_t n_pt,
                        const p2c_prim* ees, int64_t n_ee, int64_t pt_pairs, int64_t ee_pairs,
                        const p2c_queue_config* config, const p2c_refine_config* refine, void* workspace,
                        size_t workspace_bytes, p2c_certificate* certs, p2c_audit_row* rows,
                        p2c_queue_entry* children, int64_t child_cap, uint64_t* status, void* stream) {
  using namespace p2c;
  if (int rc = check_queue_args(config, refine)) return rc;
  if (n < 0 || n_pt < 0 || n_ee < 0 || pt_pairs < 0 || ee_pairs < 0 || child_cap < 0)
    return set_error(P2C_EINVAL, "exact work queue sizes must be non-negative");
  if (n > int64_t(0xFFFFFFFEu)) return set_error(P2C_EINVAL, "exact work queue holds at most 2^32 - 2 entries");
  if (status == nullptr) return set_error(P2C_EINVAL, "exact work queue status buffer is required");
  if (n > 0 && (entries == nullptr || certs == nullptr || rows == nullptr))
    return set_error(P2C_EINVAL, "exact work queue entry, certificate and row buffers are required");
  if ((n_pt > 0 && pts == nullptr) || (n_ee > 0 && ees == nullptr))
    return set_error(P2C_EINVAL, "exact work queue primitive buffers are required");
  if (refine != nullptr && child_cap > 0 && children == nullptr)
    return set_error(P2C_EINVAL, "exact work queue child buffer is required");
  const QueueLayout l = queue_layout(n, pt_pairs, ee_pairs);
  if (workspace == nullptr || workspace_bytes < l.total)
    return set_error(P2C_ECAPACITY, "exact work queue workspace is smaller than p2c_exact_queue_workspace_bytes");
  hipStream_t s = as_stream(stream);
  char* w = static_cast<char*>(workspace);
  auto* cfg = reinterpret_cast<RefCfg*>(w + l.cfg);
  auto* cnt_pt = reinterpret_cast<uint32_t*>(w + l.cnt_pt);
  auto* cnt_ee = reinterpret_cast<uint32_t*>(w + l.cnt_ee);
  auto* cnt_child = reinterpret_cast<uint32_t*>(w + l.cnt_child);
  auto* off_pt = reinterpret_cast<uint64_t*>(w + l.off_pt);
  auto* off_ee = reinterpret_cast<uint64_t*>(w + l.off_ee);
  auto* off_child = reinterpret_cast<uint64_t*>(w + l.off_child);
  auto* part = reinterpret_cast<uint64_t*>(w + l.part);
  auto* table = reinterpret_cast<uint32_t*>(w + l.table);
  auto* res_pt = reinterpret_cast<p2c_prim_result*>(w + l.res_pt);
  auto* res_ee = reinterpret_cast<p2c_prim_result*>(w + l.res_ee);
  const int64_t slots = table_slots(n);
  hipLaunchKernelGGL(queue_init_kernel, dim3(grid_for(slots, kQueueBlock)), dim3(kQueueBlock), 0, s, status, table,
                     slots);
  if (n == 0) {
    P2C_HIP_CHECK(hipGetLastError());
    return P2C_OK;
  }
  const unsigned grid = grid_for(n, kQueueBlock);
  hipLaunchKernelGGL(queue_setup_kernel, dim3(grid), dim3(kQueueBlock), 0, s, entries, n, n_pt, n_ee, cfg, cnt_pt,
                     cnt_ee, table, slots, status);
  if (int rc = exclusive_scan_u32(cnt_pt, n, off_pt, part, s)) return rc;
  if (int rc = exclusive_scan_u32(cnt_ee, n, off_ee, part, s)) return rc;
  const int mode = (config->walk_flags & P2C_CERT_CUDA_TWIN)     ? kModeCudaTwin
                   : (config->walk_flags & P2C_CERT_PYTHON_TWIN) ? kModePython
                                                                 : kModeEngine;
  if (pt_pairs > 0)
    launch_walk(P2C_FAMILY_POINT_TRIANGLE, mode, QueueSrc{pts, entries, cfg, off_pt, n, pt_pairs, true}, pt_pairs,
                res_pt, nullptr, s);
  if (ee_pairs > 0)
    launch_walk(P2C_FAMILY_EDGE_EDGE, mode, QueueSrc{ees, entries, cfg, off_ee, n, ee_pairs, false}, ee_pairs,
                res_ee, nullptr, s);
  const p2c_refine_config rc = refine ? *refine : p2c_refine_config{1, 1.0, 0, 0};
  hipLaunchKernelGGL(queue_aggregate_kernel, dim3(grid), dim3(kQueueBlock), 0, s, entries, n, res_pt, res_ee,
                     cnt_pt, cnt_ee, off_pt, off_ee, pt_pairs, ee_pairs, *config, refine ? 1 : 0, rc, certs, rows,
                     refine ? cnt_child : nullptr, status);
  if (refine) {
    if (int e = exclusive_scan_u32(cnt_child, n, off_child, part, s)) return e;
    hipLaunchKernelGGL(queue_children_kernel, dim3(grid), dim3(kQueueBlock), 0, s, entries, n, certs, cnt_child,
                       off_child, rc, children, child_cap, status);
  }
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int p2c_exact_queue_error(const uint64_t* st) {
  using namespace p2c;
  if (st == nullptr) return set_error(P2C_EINVAL, "exact work queue status is required");
  if (st[3] != 0)
    return set_error(P2C_ECAPACITY, "exact work queue pair bounds are smaller than the entries' primitive counts");
  if (st[0] != kNone) {
    const uint32_t code = uint32_t(st[0] % 64);
    return set_error(code == kErrDepth ? P2C_EUNSUPPORTED : P2C_EINVAL, queue_message(code));
  }
  // ValidateExactWorkQueueCoverage (certificate_engine.cpp:553-591): with one
  // certificate per entry, only a repeated key can fail it
  if (st[2] != kNone) return set_error(P2C_EINVAL, "exact work queue contains duplicate work_item_id for a query");
  return P2C_OK;
}

int p2c_exact_queue_host(const p2c_queue_entry* entries, int64_t n, const p2c_prim* pts, int64_t n_pt,
                         const p2c_prim* ees, int64_t n_ee, const p2c_queue_config* config,
                         const p2c_refine_config* refine, p2c_certificate* certs, p2c_audit_row* rows,
                         p2c_queue_entry* children, int64_t child_cap, int64_t* n_children) {
  using namespace p2c;
  if (int rc = check_queue_args(config, refine)) return rc;
  if (n < 0 || n_pt < 0 || n_ee < 0 || child_cap < 0)
    return set_error(P2C_EINVAL, "exact work queue sizes must be non-negative");
  if (n > 0 && entries == nullptr) return set_error(P2C_EINVAL, "exact work queue entries are required");
  // pair bounds: the counts of the families each entry may certify
  int64_t pt_pairs = 0, ee_pairs = 0;
  for (int64_t i = 0; i < n; ++i) {
    if ((entries[i].feature_family_mask & 1u) && entries[i].pt_count > 0) pt_pairs += entries[i].pt_count;
    if ((entries[i].feature_family_mask & 2u) && entries[i].ee_count > 0) ee_pairs += entries[i].ee_count;
  }
  const size_t ws = p2c_exact_queue_workspace_bytes(n, pt_pairs, ee_pairs);
  const size_t row_count = size_t(n) * (config->emit_dequeue_events ? 2 : 1);
  const size_t b_entries = sizeof(p2c_queue_entry) * size_t(n), b_pts = sizeof(p2c_prim) * size_t(n_pt),
               b_ees = sizeof(p2c_prim) * size_t(n_ee), b_certs = sizeof(p2c_certificate) * size_t(n),
               b_rows = sizeof(p2c_audit_row) * row_count,
               b_children = sizeof(p2c_queue_entry) * size_t(refine ? child_cap : 0);
  const size_t o_entries = 0, o_pts = align_up(o_entries + b_entries), o_ees = align_up(o_pts + b_pts),
               o_certs = align_up(o_ees + b_ees), o_rows = align_up(o_certs + b_certs),
               o_children = align_up(o_rows + b_rows), o_status = align_up(o_children + b_children),
               o_ws = align_up(o_status + sizeof(uint64_t) * P2C_QUEUE_STATUS_WORDS), total = o_ws + ws;
  char* d = nullptr;
  P2C_HIP_CHECK(hipMalloc(&d, total));
  int rc = P2C_OK;
  uint64_t st[P2C_QUEUE_STATUS_WORDS] = {};
  auto step = [&](hipError_t e, const char* what) {
    if (rc == P2C_OK && e != hipSuccess) rc = hip_fail(e, what);
  };
  if (b_entries) step(hipMemcpy(d + o_entries, entries, b_entries, hipMemcpyHostToDevice), "hipMemcpy entries");
  if (b_pts) step(hipMemcpy(d + o_pts, pts, b_pts, hipMemcpyHostToDevice), "hipMemcpy point-triangle primitives");
  if (b_ees) step(hipMemcpy(d + o_ees, ees, b_ees, hipMemcpyHostToDevice), "hipMemcpy edge-edge primitives");
  if (rc == P2C_OK)
    rc = p2c_exact_queue_run(reinterpret_cast<p2c_queue_entry*>(d + o_entries), n,
                             reinterpret_cast<p2c_prim*>(d + o_pts), n_pt, reinterpret_cast<p2c_prim*>(d + o_ees),
                             n_ee, pt_pairs, ee_pairs, config, refine, d + o_ws, ws,
                             reinterpret_cast<p2c_certificate*>(d + o_certs),
                             reinterpret_cast<p2c_audit_row*>(d + o_rows),
                             reinterpret_cast<p2c_queue_entry*>(d + o_children), refine ? child_cap : 0,
                             reinterpret_cast<uint64_t*>(d + o_status), nullptr);
  if (rc == P2C_OK) step(hipMemcpy(st, d + o_status, sizeof(st), hipMemcpyDeviceToHost), "hipMemcpy status");
  if (rc == P2C_OK) rc = p2c_exact_queue_error(st);
  if (rc == P2C_OK && b_certs) step(hipMemcpy(certs, d + o_certs, b_certs, hipMemcpyDeviceToHost), "hipMemcpy certificates");
  if (rc == P2C_OK && b_rows) step(hipMemcpy(rows, d + o_rows, b_rows, hipMemcpyDeviceToHost), "hipMemcpy audit rows");
  if (rc == P2C_OK && refine) {
    const int64_t got = int64_t(st[1]) < child_cap ? int64_t(st[1]) : child_cap;
    if (got > 0)
      step(hipMemcpy(children, d + o_children, sizeof(p2c_queue_entry) * size_t(got), hipMemcpyDeviceToHost),
           "hipMemcpy children");
    if (n_children) *n_children = int64_t(st[1]);
  } else if (rc == P2C_OK && n_children) {
    *n_children = 0;
  }
  (void)hipFree(d);
  return rc;
}

}  // extern "C"
