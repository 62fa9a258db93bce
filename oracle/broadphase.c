/*
 * Oracle (test infrastructure only): the reference's VF/EE candidate-set
 * definition, restated from src/tools/scene_object_envelope_native_ti_walltime.cpp.
 *
 *   swept boxes   Extend/Expand/SweptVertexBox/SweptFaceBox/SweptEdgeBox :256-302
 *   sweep         SweepPairs :603-649 (stable sort by min.x, active list,
 *                 inclusive FP64 overlap :270-274)
 *   VF set        BuildVFCandidates :651-671: (vertex, face) overlapping, object
 *                 pair allowed, not (same object and face contains vertex)
 *   EE set        BuildEECandidates :673-696: (edge i < edge j) overlapping,
 *                 allowed, not (same object and edges share a vertex)
 * plus an O(n^2) brute force with the same rules (what the reference's
 * GenerateRawCandidatesCpu does for proxies, candidate_buffer.cpp:144-181).
 *
 * Boxes are 6 doubles: min x, y, z, max x, y, z.  `allowed` is the object-pair
 * byte matrix of PrepareScene :541-550.  Outputs are (a, b) int32 pairs; the
 * return value is the full count even when it exceeds `cap`.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static void box_init(double* b) {
  b[0] = b[1] = b[2] = INFINITY;
  b[3] = b[4] = b[5] = -INFINITY;
}
/* Extend with std::min / std::max semantics */
static void box_extend(double* b, const double* p) {
  for (int k = 0; k < 3; ++k) {
    b[k] = (p[k] < b[k]) ? p[k] : b[k];
    b[3 + k] = (b[3 + k] < p[k]) ? p[k] : b[3 + k];
  }
}
static void box_expand(double* b, double d) {
  for (int k = 0; k < 3; ++k) {
    b[k] -= d;
    b[3 + k] += d;
  }
}
static int overlap(const double* a, const double* b) {
  return a[0] <= b[3] && b[0] <= a[3] && a[1] <= b[4] && b[1] <= a[4] && a[2] <= b[5] && b[2] <= a[5];
}

void orc_swept_boxes(const double* x0, const double* x1, int64_t nv, const int32_t* faces, int64_t nf,
                     const int32_t* edges, int64_t ne, double thickness, double* vbox, double* fbox,
                     double* ebox) {
  for (int64_t v = 0; v < nv; ++v) {
    double* b = vbox + v * 6;
    box_init(b);
    box_extend(b, x0 + v * 3);
    box_extend(b, x1 + v * 3);
    box_expand(b, thickness);
  }
  for (int64_t f = 0; f < nf; ++f) {
    double* b = fbox + f * 6;
    box_init(b);
    for (int k = 0; k < 3; ++k) {
      box_extend(b, x0 + (int64_t)faces[f * 3 + k] * 3);
      box_extend(b, x1 + (int64_t)faces[f * 3 + k] * 3);
    }
    box_expand(b, thickness);
  }
  for (int64_t e = 0; e < ne; ++e) {
    double* b = ebox + e * 6;
    const int64_t a = edges[e * 2], c = edges[e * 2 + 1];
    box_init(b);
    box_extend(b, x0 + a * 3);
    box_extend(b, x0 + c * 3);
    box_extend(b, x1 + a * 3);
    box_extend(b, x1 + c * 3);
    box_expand(b, thickness);
  }
}

static int pair_allowed(const uint8_t* allowed, int32_t nobj, int32_t a, int32_t b) {
  if (a < 0 || b < 0 || a >= nobj || b >= nobj) return 0;
  return allowed[(int64_t)a * nobj + b] != 0;
}

typedef struct {
  const double* box;
  int64_t n;
} Sorted;

/* stable sort of indices by min.x (merge sort keeps ties in index order) */
static void stable_sort_by_minx(int32_t* idx, int64_t n, const double* box) {
  int32_t* tmp = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t w = 1; w < n; w *= 2) {
    for (int64_t lo = 0; lo < n; lo += 2 * w) {
      int64_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
      int64_t i = lo, j = mid, k = lo;
      while (i < mid && j < hi) {
        /* take right only when strictly smaller: stable */
        if (box[(int64_t)idx[j] * 6] < box[(int64_t)idx[i] * 6]) tmp[k++] = idx[j++];
        else tmp[k++] = idx[i++];
      }
      while (i < mid) tmp[k++] = idx[i++];
      while (j < hi) tmp[k++] = idx[j++];
    }
    memcpy(idx, tmp, sizeof(int32_t) * (size_t)n);
  }
  free(tmp);
}

typedef int (*accept_fn)(const void* ctx, int32_t a, int32_t b);

/* SweepPairs: lhs items (boxes lbox, indices 0..nl-1) against rhs items. */
static int64_t sweep(const double* lbox, int64_t nl, const double* rbox, int64_t nr, accept_fn fn, const void* ctx,
                     int32_t* out, int64_t cap) {
  int32_t* lo = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nl > 0 ? nl : 1));
  int32_t* ro = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nr > 0 ? nr : 1));
  int32_t* active = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nr > 0 ? nr : 1));
  for (int64_t i = 0; i < nl; ++i) lo[i] = (int32_t)i;
  for (int64_t i = 0; i < nr; ++i) ro[i] = (int32_t)i;
  stable_sort_by_minx(lo, nl, lbox);
  stable_sort_by_minx(ro, nr, rbox);
  int64_t nact = 0, cursor = 0, count = 0;
  for (int64_t li = 0; li < nl; ++li) {
    const int32_t ai = lo[li];
    const double* ab = lbox + (int64_t)ai * 6;
    while (cursor < nr && rbox[(int64_t)ro[cursor] * 6] <= ab[3]) active[nact++] = ro[cursor++];
    int64_t w = 0;
    for (int64_t k = 0; k < nact; ++k)
      if (rbox[(int64_t)active[k] * 6 + 3] >= ab[0]) active[w++] = active[k];
    nact = w;
    for (int64_t k = 0; k < nact; ++k) {
      const int32_t bi = active[k];
      if (!overlap(ab, rbox + (int64_t)bi * 6)) continue;
      if (fn(ctx, ai, bi)) {
        if (count < cap) {
          out[count * 2] = ai;
          out[count * 2 + 1] = bi;
        }
        ++count;
      }
    }
  }
  free(lo);
  free(ro);
  free(active);
  return count;
}

typedef struct {
  const int32_t* vid;  /* vertex proxy slot -> vertex id */
  const int32_t* vobj; /* by vertex id */
  const int32_t* faces;
  const int32_t* fobj;
  const int32_t* edges;
  const int32_t* eobj;
  const uint8_t* allowed;
  int32_t nobj;
} Rules;

static int vf_accept(const void* c, int32_t slot, int32_t f) {
  const Rules* r = (const Rules*)c;
  const int32_t v = r->vid[slot];
  const int32_t ov = r->vobj[v], of = r->fobj[f];
  if (!pair_allowed(r->allowed, r->nobj, ov, of)) return 0;
  const int32_t* fv = r->faces + (int64_t)f * 3;
  if (ov == of && (fv[0] == v || fv[1] == v || fv[2] == v)) return 0;
  return 1;
}

static int ee_accept(const void* c, int32_t a, int32_t b) {
  const Rules* r = (const Rules*)c;
  if (a >= b) return 0;
  const int32_t oa = r->eobj[a], ob = r->eobj[b];
  if (!pair_allowed(r->allowed, r->nobj, oa, ob)) return 0;
  const int32_t* ea = r->edges + (int64_t)a * 2;
  const int32_t* eb = r->edges + (int64_t)b * 2;
  if (oa == ob && (ea[0] == eb[0] || ea[0] == eb[1] || ea[1] == eb[0] || ea[1] == eb[1])) return 0;
  return 1;
}

/* Vertex proxies exist only for vertices that belong to an object (:552-559). */
static int64_t vertex_proxies(const double* vbox, const int32_t* vobj, int64_t nv, double** pbox, int32_t** pid) {
  int64_t n = 0;
  for (int64_t v = 0; v < nv; ++v) n += vobj[v] >= 0;
  *pbox = (double*)malloc(sizeof(double) * 6 * (size_t)(n > 0 ? n : 1));
  *pid = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int64_t k = 0;
  for (int64_t v = 0; v < nv; ++v)
    if (vobj[v] >= 0) {
      memcpy(*pbox + k * 6, vbox + v * 6, sizeof(double) * 6);
      (*pid)[k++] = (int32_t)v;
    }
  return n;
}

int64_t orc_sweep_vf(const double* vbox, const int32_t* vobj, int64_t nv, const double* fbox, const int32_t* faces,
                     const int32_t* fobj, int64_t nf, const uint8_t* allowed, int32_t nobj, int32_t* out,
                     int64_t cap) {
  double* pbox;
  int32_t* pid;
  const int64_t np = vertex_proxies(vbox, vobj, nv, &pbox, &pid);
  Rules r = {pid, vobj, faces, fobj, NULL, NULL, allowed, nobj};
  const int64_t n = sweep(pbox, np, fbox, nf, vf_accept, &r, out, cap);
  /* report vertex ids, not proxy slots */
  const int64_t m = n < cap ? n : cap;
  for (int64_t i = 0; i < m; ++i) out[i * 2] = pid[out[i * 2]];
  free(pbox);
  free(pid);
  return n;
}

int64_t orc_sweep_ee(const double* ebox, const int32_t* edges, const int32_t* eobj, int64_t ne,
                     const uint8_t* allowed, int32_t nobj, int32_t* out, int64_t cap) {
  Rules r = {NULL, NULL, NULL, NULL, edges, eobj, allowed, nobj};
  return sweep(ebox, ne, ebox, ne, ee_accept, &r, out, cap);
}

int64_t orc_brute_vf(const double* vbox, const int32_t* vobj, int64_t nv, const double* fbox, const int32_t* faces,
                     const int32_t* fobj, int64_t nf, const uint8_t* allowed, int32_t nobj, int32_t* out,
                     int64_t cap) {
  int64_t count = 0;
  for (int64_t v = 0; v < nv; ++v) {
    if (vobj[v] < 0) continue;
    for (int64_t f = 0; f < nf; ++f) {
      if (!overlap(vbox + v * 6, fbox + f * 6)) continue;
      const int32_t ov = vobj[v], of = fobj[f];
      if (!pair_allowed(allowed, nobj, ov, of)) continue;
      const int32_t* fv = faces + f * 3;
      if (ov == of && (fv[0] == v || fv[1] == v || fv[2] == v)) continue;
      if (count < cap) {
        out[count * 2] = (int32_t)v;
        out[count * 2 + 1] = (int32_t)f;
      }
      ++count;
    }
  }
  return count;
}

int64_t orc_brute_ee(const double* ebox, const int32_t* edges, const int32_t* eobj, int64_t ne,
                     const uint8_t* allowed, int32_t nobj, int32_t* out, int64_t cap) {
  int64_t count = 0;
  for (int64_t a = 0; a < ne; ++a)
    for (int64_t b = a + 1; b < ne; ++b) {
      if (!overlap(ebox + a * 6, ebox + b * 6)) continue;
      const int32_t oa = eobj[a], ob = eobj[b];
      if (!pair_allowed(allowed, nobj, oa, ob)) continue;
      const int32_t* ea = edges + a * 2;
      const int32_t* eb = edges + b * 2;
      if (oa == ob && (ea[0] == eb[0] || ea[0] == eb[1] || ea[1] == eb[0] || ea[1] == eb[1])) continue;
      if (count < cap) {
        out[count * 2] = (int32_t)a;
        out[count * 2 + 1] = (int32_t)b;
      }
      ++count;
    }
  return count;
}
