/*
 * Oracle (test infrastructure only): TinySTPF candidate scoring restated in
 * C from src/tools/scene_object_envelope_native_ti_walltime.cpp:
 *   Mix64 / UnitHash            :773-782
 *   FeatureFromCandidate        :893-984, with the index slot of :986-993
 *   PredictTinyStpf             :1068-1091
 *   Sigmoid / ScoreCandidate    :1093-1131
 * FP64, no contraction (oracle/Makefile), sequential sums in the reference's
 * order.  Pinned by tests/test_oracle_mlp.py: hand-derived feature values on
 * simple geometry, torch float64 as the library oracle of the MLP on the
 * bundled frozen weights, and planted mistakes.
 */
#include <math.h>
#include <stdint.h>

#include "oracle.h"

int orc_mlp_mutation = 0; /* planted mistakes: 1 transposes w1, 2 drops the second ReLU */

uint64_t orc_mix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}

double orc_unit_hash(uint64_t x) { return (double)(orc_mix64(x) >> 11) * (1.0 / 9007199254740992.0); }

static double dmin(double a, double b) { return b < a ? b : a; } /* std::min */
static double dmax(double a, double b) { return a < b ? b : a; } /* std::max */

typedef struct {
  double x, y, z;
} v3;
static v3 add3(v3 a, v3 b) { return (v3){a.x + b.x, a.y + b.y, a.z + b.z}; }
static v3 sub3(v3 a, v3 b) { return (v3){a.x - b.x, a.y - b.y, a.z - b.z}; }
static v3 scale3(v3 a, double s) { return (v3){a.x * s, a.y * s, a.z * s}; }
static double dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static v3 cross3(v3 a, v3 b) { return (v3){a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
static double norm3(v3 a) { return sqrt(dot3(a, a)); }

void orc_tiny_features(int is_vf, const double* rows24, uint64_t index, double f[32]) {
  v3 r[8];
  for (int i = 0; i < 8; ++i) r[i] = (v3){rows24[3 * i], rows24[3 * i + 1], rows24[3 * i + 2]};
  v3 mn = {INFINITY, INFINITY, INFINITY}, mx = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = 0; i < 8; ++i) {
    mn = (v3){dmin(mn.x, r[i].x), dmin(mn.y, r[i].y), dmin(mn.z, r[i].z)};
    mx = (v3){dmax(mx.x, r[i].x), dmax(mx.y, r[i].y), dmax(mx.z, r[i].z)};
  }
  const v3 extent = sub3(mx, mn);
  v3 motion[4];
  for (int i = 0; i < 4; ++i) motion[i] = sub3(r[4 + i], r[i]);
  const v3 ma = scale3(add3(motion[0], motion[1]), 0.5);
  const v3 mb = scale3(add3(motion[2], motion[3]), 0.5);
  const double rel = norm3(sub3(ma, mb));
  double speed = 0.0;
  for (int i = 0; i < 4; ++i) speed = dmax(speed, norm3(motion[i]));
  double gap0 = INFINITY, gap1 = INFINITY;
  for (int i = 0; i < 4; ++i)
    for (int j = i + 1; j < 4; ++j) gap0 = dmin(gap0, norm3(sub3(r[i], r[j])));
  for (int i = 0; i < 4; ++i)
    for (int j = i + 1; j < 4; ++j) gap1 = dmin(gap1, norm3(sub3(r[4 + i], r[4 + j])));
  v3 mean0 = {0, 0, 0}, mean1 = {0, 0, 0}, mean_all = {0, 0, 0};
  for (int i = 0; i < 4; ++i) {
    mean0 = add3(mean0, r[i]);
    mean1 = add3(mean1, r[4 + i]);
  }
  mean0 = scale3(mean0, 0.25);
  mean1 = scale3(mean1, 0.25);
  for (int i = 0; i < 8; ++i) mean_all = add3(mean_all, r[i]);
  mean_all = scale3(mean_all, 0.125);
  double variance = 0.0;
  for (int i = 0; i < 8; ++i) {
    const v3 d = sub3(r[i], mean_all);
    variance += dot3(d, d);
  }
  variance /= 8.0;
  f[0] = is_vf ? 1.0 : 0.0;
  f[1] = is_vf ? 0.0 : 1.0;
  f[2] = extent.x;
  f[3] = extent.y;
  f[4] = extent.z;
  f[5] = norm3(extent);
  f[6] = log1p(dmax(extent.x, 1.0e-9) * dmax(extent.y, 1.0e-9) * dmax(extent.z, 1.0e-9));
  f[7] = rel;
  f[8] = speed;
  f[9] = gap0;
  f[10] = gap1;
  f[11] = dmin(gap0, gap1);
  f[12] = mean_all.x;
  f[13] = mean_all.y;
  f[14] = mean_all.z;
  f[15] = norm3(sub3(mean1, mean0));
  f[16] = 0.0;
  for (int i = 0; i < 8; ++i) f[16] = dmax(f[16], dmax(dmax(fabs(r[i].x), fabs(r[i].y)), fabs(r[i].z)));
  f[17] = sqrt(variance);
  for (int i = 0; i < 4; ++i) f[18 + i] = norm3(motion[i]);
  for (int i = 0; i < 3; ++i) f[22 + i] = norm3(sub3(r[1 + i], r[0]));
  f[25] = 1.0 / (1.0 + dmin(gap0, gap1));
  f[26] = f[8] / (1.0 + f[5]);
  f[27] = dot3(motion[0], motion[3]);
  f[28] = norm3(cross3(motion[0], motion[3]));
  f[29] = norm3(sub3(mean0, mean1));
  f[30] = log1p((double)index);
  f[31] = 1.0;
}

/* packed weights: mean[32] std[32] w0[64][32] b0[64] w1[32][64] b1[32] w2[32] b2 */
double orc_tiny_predict(const double* w, const double f[32]) {
  const double *mean = w, *sd = w + 32, *w0 = w + 64, *b0 = w0 + 64 * 32, *w1 = b0 + 64, *b1 = w1 + 32 * 64,
               *w2 = b1 + 32, b2 = w2[32];
  double h0[64], h1[32];
  for (int o = 0; o < 64; ++o) {
    double sum = b0[o];
    for (int i = 0; i < 32; ++i) {
      const double x = (f[i] - mean[i]) / dmax(sd[i], 1.0e-12);
      sum += w0[o * 32 + i] * x;
    }
    h0[o] = dmax(0.0, sum);
  }
  for (int o = 0; o < 32; ++o) {
    double sum = b1[o];
    for (int i = 0; i < 64; ++i) sum += (orc_mlp_mutation == 1 ? w1[i * 32 + o] : w1[o * 64 + i]) * h0[i];
    h1[o] = orc_mlp_mutation == 2 ? sum : dmax(0.0, sum);
  }
  double out = b2;
  for (int i = 0; i < 32; ++i) out += w2[i] * h1[i];
  return out;
}

static double sigmoid(double v) {
  if (v >= 0.0) {
    const double z = exp(-v);
    return 1.0 / (1.0 + z);
  }
  const double z = exp(v);
  return z / (1.0 + z);
}

/* policy: 0 learned, 1 learned_residual, 2 proximity, 3 motion, 4 random */
double orc_tiny_score(int policy, int is_vf, const double* rows24, uint64_t index, uint64_t group_key,
                      const double* w) {
  if (policy == 4) return orc_unit_hash(index ^ (is_vf ? 0x767666ULL : 0x656565ULL) ^ (group_key << 1));
  double f[32];
  orc_tiny_features(is_vf, rows24, index, f);
  if (policy == 0) return orc_tiny_predict(w, f);
  if (policy == 1) return f[25] + 0.08 * sigmoid(orc_tiny_predict(w, f)) + 0.004 * f[26] - 0.00025 * f[30];
  if (policy == 2) return f[25];
  return f[8];
}

void orc_tiny_score_batch(int policy, int is_vf, const double* rows, const uint64_t* index, const uint64_t* keys,
                          int64_t n, const double* w, double* out) {
  for (int64_t c = 0; c < n; ++c) out[c] = orc_tiny_score(policy, is_vf, rows + 24 * c, index[c], keys[c], w);
}
