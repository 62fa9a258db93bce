/*
 * Oracle (test infrastructure only): proxy-scene candidates restated in C.
 *   GenerateRawCandidatesCpu   src/cpp/rt_candidate/candidate_buffer.cpp:144-181
 *     pairs i < j: different objects, same slab, strict time overlap
 *     (:27-29), inclusive AABB overlap (:18-25); PairKey (:31-35);
 *     MergeMotionBound (:37-46, max then float); flags valid | overlap
 *   CompactRawCandidates       :183-228 with MakeCandidateKey (:104-124):
 *     key (query, slab, first object, first patch, second object, second
 *     patch, first type, second type), first = the smaller (object, patch,
 *     proxy id); duplicates add rt_hit_count, max the motion bound, OR the
 *     flags; candidate_id = 1.. in key order
 *   ComputeCandidateDensityStats src/cpp/rt_candidate/candidate_stats.cpp:16-43, 128-170
 * Validation messages follow proxy_scene.cpp:18-154.  Pinned by golden vectors
 * from the reference compiled in place (oracle/ref_proxy_shim.cpp,
 * tests/golden/make_proxy_golden.py) including the reference's KAT scene
 * (src/tests/cpp/test_rt_candidate.cpp:97-150), and by brute-force set checks
 * (tests/test_oracle_proxy.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int aabb_overlap(const double* a, const double* b) {
  for (int k = 0; k < 3; ++k)
    if (a[3 + k] < b[k] || b[3 + k] < a[k]) return 0;
  return 1;
}

static float fmaxf_std(double a, double b) { return (float)(a < b ? b : a); } /* std::max then cast */

typedef struct {
  uint32_t slab, obj_a, patch_a, obj_b, patch_b, type_a, type_b;
  int64_t hit; /* index of the raw hit */
} key_t_;

static int key_cmp(const void* pa, const void* pb) {
  const key_t_* a = (const key_t_*)pa;
  const key_t_* b = (const key_t_*)pb;
  const uint32_t fa[7] = {a->slab, a->obj_a, a->patch_a, a->obj_b, a->patch_b, a->type_a, a->type_b};
  const uint32_t fb[7] = {b->slab, b->obj_a, b->patch_a, b->obj_b, b->patch_b, b->type_a, b->type_b};
  for (int i = 0; i < 7; ++i)
    if (fa[i] != fb[i]) return fa[i] < fb[i] ? -1 : 1;
  return a->hit < b->hit ? -1 : (a->hit > b->hit);
}

int orc_proxy_generate(uint64_t query_id, const orc_proxy_set* s, int64_t raw_cap, orc_raw_hit* raw, int64_t* n_raw,
                       int64_t cand_cap, orc_candidate* cand, int64_t* n_cand, orc_density* d) {
  if (query_id == 0) return -10;
  const int64_t n = s->n;
  for (int64_t i = 0; i < n; ++i) {
    if (s->proxy_type[i] != 1 && s->proxy_type[i] != 2) return -11;
    if (!isfinite(s->t0[i]) || !isfinite(s->t1[i]) || s->t0[i] >= s->t1[i]) return -12;
    for (int k = 0; k < 3; ++k) {
      if (!isfinite(s->bounds[6 * i + k]) || !isfinite(s->bounds[6 * i + 3 + k])) return -13;
      if (s->bounds[6 * i + k] > s->bounds[6 * i + 3 + k]) return -14;
    }
  }
  /* raw hits, (i, j) in loop order */
  int64_t m = 0;
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = i + 1; j < n; ++j) {
      if (s->object_id[i] == s->object_id[j] || s->slab_id[i] != s->slab_id[j]) continue;
      if (!(s->t0[i] < s->t1[j] && s->t0[j] < s->t1[i])) continue;
      if (!aabb_overlap(s->bounds + 6 * i, s->bounds + 6 * j)) continue;
      if (m < raw_cap) {
        orc_raw_hit* h = &raw[m];
        memset(h, 0, sizeof(*h));
        h->query_id = query_id;
        h->pair_key = ((uint64_t)(uint32_t)i << 32) | (uint32_t)j;
        h->proxy_a_index = (uint32_t)i;
        h->proxy_b_index = (uint32_t)j;
        h->slab_id = s->slab_id[i];
        h->rt_hit_count = 1;
        h->flags = 3u;
        for (int k = 0; k < 4; ++k) h->motion_bound[k] = fmaxf_std(s->motion[4 * i + k], s->motion[4 * j + k]);
      }
      ++m;
    }
  *n_raw = m;
  if (m > raw_cap) return -20;
  /* compaction */
  key_t_* keys = (key_t_*)malloc(sizeof(key_t_) * (size_t)(m > 0 ? m : 1));
  if (!keys) return -21;
  for (int64_t h = 0; h < m; ++h) {
    uint32_t a = raw[h].proxy_a_index, b = raw[h].proxy_b_index;
    /* first = the smaller (object, patch, proxy id) */
    int swap = 0;
    if (s->object_id[a] != s->object_id[b]) swap = s->object_id[a] > s->object_id[b];
    else if (s->patch_id[a] != s->patch_id[b]) swap = s->patch_id[a] > s->patch_id[b];
    else swap = a > b;
    if (swap) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    keys[h] = (key_t_){raw[h].slab_id, s->object_id[a], s->patch_id[a], s->object_id[b], s->patch_id[b],
                       s->proxy_type[a], s->proxy_type[b], h};
  }
  qsort(keys, (size_t)m, sizeof(key_t_), key_cmp);
  int64_t c = 0;
  uint64_t rt_sum = 0;
  for (int64_t k = 0; k < m; ++k) {
    const orc_raw_hit* h = &raw[keys[k].hit];
    const int same = k > 0 && key_cmp(&(key_t_){keys[k].slab, keys[k].obj_a, keys[k].patch_a, keys[k].obj_b,
                                                 keys[k].patch_b, keys[k].type_a, keys[k].type_b, 0},
                                      &(key_t_){keys[k - 1].slab, keys[k - 1].obj_a, keys[k - 1].patch_a,
                                                 keys[k - 1].obj_b, keys[k - 1].patch_b, keys[k - 1].type_a,
                                                 keys[k - 1].type_b, 0}) == 0;
    if (!same) {
      if (c < cand_cap) {
        orc_candidate* r = &cand[c];
        memset(r, 0, sizeof(*r));
        r->candidate_id = (uint64_t)(c + 1);
        r->query_id = query_id;
        r->slab_id = keys[k].slab;
        r->object_a_id = keys[k].obj_a;
        r->patch_a_id = keys[k].patch_a;
        r->object_b_id = keys[k].obj_b;
        r->patch_b_id = keys[k].patch_b;
        r->proxy_type_a = keys[k].type_a;
        r->proxy_type_b = keys[k].type_b;
        r->rt_hit_count = h->rt_hit_count;
        r->flags = h->flags;
        for (int q = 0; q < 4; ++q) r->motion_bound[q] = h->motion_bound[q];
      }
      ++c;
    } else if (c - 1 < cand_cap) {
      orc_candidate* r = &cand[c - 1];
      r->rt_hit_count += h->rt_hit_count;
      for (int q = 0; q < 4; ++q) r->motion_bound[q] = r->motion_bound[q] < h->motion_bound[q] ? h->motion_bound[q]
                                                                                                : r->motion_bound[q];
      r->flags |= h->flags;
    }
    rt_sum += h->rt_hit_count;
  }
  free(keys);
  *n_cand = c;
  if (c > cand_cap) return -22;
  /* density */
  uint64_t objects = 0, slabs = 0, cross = 0;
  for (int64_t i = 0; i < n; ++i) {
    int new_obj = 1, new_slab = 1;
    for (int64_t j = 0; j < i; ++j) {
      if (s->object_id[j] == s->object_id[i]) new_obj = 0;
      if (s->slab_id[j] == s->slab_id[i]) new_slab = 0;
    }
    objects += (uint64_t)new_obj;
    slabs += (uint64_t)new_slab;
    for (int64_t j = i + 1; j < n; ++j) cross += (s->object_id[i] != s->object_id[j] && s->slab_id[i] == s->slab_id[j]);
  }
#define RATIO(a, b) ((b) == 0 ? 0.0 : (double)(a) / (double)(b))
  d->proxy_count = (uint64_t)n;
  d->object_count = objects;
  d->slab_count = slabs;
  d->cross_object_same_slab_pair_count = cross;
  d->raw_hit_count = (uint64_t)m;
  d->compact_candidate_count = (uint64_t)c;
  d->raw_hits_per_proxy = RATIO(m, n);
  d->candidates_per_proxy = RATIO(c, n);
  d->candidates_per_slab = RATIO(c, slabs);
  d->aabb_overlap_ratio = RATIO(m, cross);
  d->avg_rt_hits_per_candidate = RATIO(rt_sum, c);
#undef RATIO
  return 0;
}
