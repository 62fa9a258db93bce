// Scorers of the scheduling path on gfx950.  They only order certificate
// work; every answer is still decided by an exact certificate.
//
// 1. TinySTPF candidate scores, FP64, one lane per candidate
//    (src/tools/scene_object_envelope_native_ti_walltime.cpp):
//      GroupKey :756-771   object pair (min, max) packed in a u64
//      UnitHash :773-782   SplitMix64 finaliser, top 53 bits as a double
//      features :893-993   32 per candidate (extent, motion, gaps, spread,
//                          log1p(index) in slot 30)
//      predict  :1068-1091 z-score (std floor 1e-12), 32-64 ReLU, 64-32 ReLU,
//                          32-1; summation orders kept, no contraction
//      policies :1102-1131 learned, learned_residual, proximity, motion, random
//    The weights (4289 doubles, layout in p2c_hip.h) are read with
//    wave-uniform indices, so they come through the scalar cache.
// 2. STPF priority trunk, FP32 (src/python/p2cccd/proposal/stpf_model.py:
//    45-93): 32 -> 128 GELU -> 128 GELU -> 1, sigmoid, one lane per row with
//    the first hidden layer in registers and the second layer folded into the
//    output as it is produced.
#include <cmath>

#include "p2c_hip.h"
#include "runtime.h"

namespace p2c {
namespace {

constexpr int kBlock = 256;

struct V3 {
  double x, y, z;
};
__device__ inline V3 vadd(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline V3 vscale(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ inline double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline V3 vcross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ inline double vnorm(V3 a) { return __builtin_sqrt(vdot(a, a)); }
// std::min / std::max semantics (first argument kept unless the second is strictly better)
__device__ inline double smin(double a, double b) { return b < a ? b : a; }
__device__ inline double smax(double a, double b) { return a < b ? b : a; }

__device__ inline V3 load3(const double* x, int32_t i) { return {x[3 * i], x[3 * i + 1], x[3 * i + 2]}; }

__device__ inline uint64_t mix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}
__device__ inline double unit_hash(uint64_t x) { return (double)(mix64(x) >> 11) * (1.0 / 9007199254740992.0); }

__device__ inline double sigmoid64(double v) {
  if (v >= 0.0) {
    const double z = exp(-v);
    return 1.0 / (1.0 + z);
  }
  const double z = exp(v);
  return z / (1.0 + z);
}

// The 32 candidate features; rows = the 4 vertices at t0 then at t1.
__device__ void tiny_features(int is_vf, const V3 r[8], uint64_t index, double f[32]) {
  V3 mn = r[0], mx = r[0];
#pragma unroll
  for (int i = 1; i < 8; ++i) {
    mn = {smin(mn.x, r[i].x), smin(mn.y, r[i].y), smin(mn.z, r[i].z)};
    mx = {smax(mx.x, r[i].x), smax(mx.y, r[i].y), smax(mx.z, r[i].z)};
  }
  const V3 extent = vsub(mx, mn);
  V3 motion[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) motion[i] = vsub(r[4 + i], r[i]);
  const V3 mean_a = vscale(vadd(motion[0], motion[1]), 0.5);
  const V3 mean_b = vscale(vadd(motion[2], motion[3]), 0.5);
  const double rel = vnorm(vsub(mean_a, mean_b));
  double speed = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) speed = smax(speed, vnorm(motion[i]));
  double gap0 = __builtin_inf(), gap1 = __builtin_inf();
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = i + 1; j < 4; ++j) {
      gap0 = smin(gap0, vnorm(vsub(r[i], r[j])));
      gap1 = smin(gap1, vnorm(vsub(r[4 + i], r[4 + j])));
    }
  V3 mean0 = {0, 0, 0}, mean1 = {0, 0, 0}, mean_all = {0, 0, 0};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    mean0 = vadd(mean0, r[i]);
    mean1 = vadd(mean1, r[4 + i]);
  }
  mean0 = vscale(mean0, 0.25);
  mean1 = vscale(mean1, 0.25);
#pragma unroll
  for (int i = 0; i < 8; ++i) mean_all = vadd(mean_all, r[i]);
  mean_all = vscale(mean_all, 0.125);
  double variance = 0.0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const V3 d = vsub(r[i], mean_all);
    variance += vdot(d, d);
  }
  variance /= 8.0;
  f[0] = is_vf ? 1.0 : 0.0;
  f[1] = is_vf ? 0.0 : 1.0;
  f[2] = extent.x;
  f[3] = extent.y;
  f[4] = extent.z;
  f[5] = vnorm(extent);
  f[6] = log1p(smax(extent.x, 1.0e-9) * smax(extent.y, 1.0e-9) * smax(extent.z, 1.0e-9));
  f[7] = rel;
  f[8] = speed;
  f[9] = gap0;
  f[10] = gap1;
  f[11] = smin(gap0, gap1);
  f[12] = mean_all.x;
  f[13] = mean_all.y;
  f[14] = mean_all.z;
  f[15] = vnorm(vsub(mean1, mean0));
  double amax = 0.0;
#pragma unroll
  for (int i = 0; i < 8; ++i) amax = smax(amax, smax(smax(fabs(r[i].x), fabs(r[i].y)), fabs(r[i].z)));
  f[16] = amax;
  f[17] = __builtin_sqrt(variance);
  f[18] = vnorm(motion[0]);
  f[19] = vnorm(motion[1]);
  f[20] = vnorm(motion[2]);
  f[21] = vnorm(motion[3]);
  f[22] = vnorm(vsub(r[1], r[0]));
  f[23] = vnorm(vsub(r[2], r[0]));
  f[24] = vnorm(vsub(r[3], r[0]));
  f[25] = 1.0 / (1.0 + smin(gap0, gap1));
  f[26] = f[8] / (1.0 + f[5]);
  f[27] = vdot(motion[0], motion[3]);
  f[28] = vnorm(vcross(motion[0], motion[3]));
  f[29] = vnorm(vsub(mean0, mean1));
  f[30] = log1p((double)index);
  f[31] = 1.0;
}

// Packed TinySTPF weights (doubles): mean[32] std[32] w0[64][32] b0[64]
// w1[32][64] b1[32] w2[32] b2.
constexpr int kMean = 0, kStd = 32, kW0 = 64, kB0 = kW0 + 64 * 32, kW1 = kB0 + 64, kB1 = kW1 + 32 * 64,
              kW2 = kB1 + 32, kB2 = kW2 + 32, kTinyWeights = kB2 + 1;
static_assert(kTinyWeights == P2C_TINY_STPF_WEIGHTS, "TinySTPF weight layout");

__device__ double tiny_predict(const double* __restrict__ w, const double f[32]) {
  double x[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) x[i] = (f[i] - w[kMean + i]) / smax(w[kStd + i], 1.0e-12);
  double h1[32];
#pragma unroll
  for (int j = 0; j < 32; ++j) h1[j] = w[kB1 + j];
  // hidden unit o of the first layer, then its term of every second-layer sum
  // (each second-layer sum still adds its terms in ascending o)
  for (int o = 0; o < 64; ++o) {
    double s = w[kB0 + o];
#pragma unroll
    for (int i = 0; i < 32; ++i) s += w[kW0 + o * 32 + i] * x[i];
    const double h0 = smax(0.0, s);
#pragma unroll
    for (int j = 0; j < 32; ++j) h1[j] += w[kW1 + j * 64 + o] * h0;
  }
  double out = w[kB2];
#pragma unroll
  for (int j = 0; j < 32; ++j) out += w[kW2 + j] * smax(0.0, h1[j]);
  return out;
}

__global__ __launch_bounds__(kBlock) void tiny_score_kernel(int is_vf, const double* __restrict__ x0,
                                                             const double* __restrict__ x1,
                                                             const int32_t* __restrict__ topo,
                                                             const int32_t* __restrict__ pairs, int64_t n_host,
                                                             const uint64_t* __restrict__ totals,
                                                             const int32_t* __restrict__ obj_a,
                                                             const int32_t* __restrict__ obj_b, int policy,
                                                             const double* __restrict__ w, double* __restrict__ scores,
                                                             uint64_t* __restrict__ keys) {
  int64_t n = n_host;
  if (totals != nullptr) {
    const int64_t t = (int64_t)totals[is_vf ? 0 : 1];
    n = t < n ? t : n;
  }
  for (int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; c < n; c += int64_t(gridDim.x) * blockDim.x) {
    const int32_t a = pairs[2 * c], b = pairs[2 * c + 1];
    int32_t vid[4];
    if (is_vf) {
      vid[0] = a;
      vid[1] = topo[3 * b];
      vid[2] = topo[3 * b + 1];
      vid[3] = topo[3 * b + 2];
    } else {
      vid[0] = topo[2 * a];
      vid[1] = topo[2 * a + 1];
      vid[2] = topo[2 * b];
      vid[3] = topo[2 * b + 1];
    }
    int32_t oa = obj_a[a], ob = obj_b[b];
    if (ob < oa) {
      const int32_t t = oa;
      oa = ob;
      ob = t;
    }
    const uint64_t key = ((uint64_t)(uint32_t)oa << 32) | (uint32_t)ob;
    if (keys != nullptr) keys[c] = key;
    if (scores == nullptr) continue;
    double s;
    if (policy == P2C_SCORE_RANDOM) {
      s = unit_hash((uint64_t)c ^ (is_vf ? 0x767666ull : 0x656565ull) ^ (key << 1));
    } else {
      V3 r[8];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        r[i] = load3(x0, vid[i]);
        r[4 + i] = load3(x1, vid[i]);
      }
      double f[32];
      tiny_features(is_vf, r, (uint64_t)c, f);
      if (policy == P2C_SCORE_LEARNED) {
        s = tiny_predict(w, f);
      } else if (policy == P2C_SCORE_LEARNED_RESIDUAL) {
        const double learned = sigmoid64(tiny_predict(w, f));
        s = f[25] + 0.08 * learned + 0.004 * f[26] - 0.00025 * f[30];
      } else if (policy == P2C_SCORE_PROXIMITY) {
        s = f[25];
      } else {
        s = f[8];
      }
    }
    scores[c] = s;
  }
}

__global__ __launch_bounds__(kBlock) void tiny_features_kernel(int is_vf, const double* __restrict__ rows,
                                                                const uint64_t* __restrict__ index, int64_t n,
                                                                double* __restrict__ out) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  V3 r[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) r[i] = {rows[24 * c + 3 * i], rows[24 * c + 3 * i + 1], rows[24 * c + 3 * i + 2]};
  double f[32];
  tiny_features(is_vf, r, index[c], f);
#pragma unroll
  for (int i = 0; i < 32; ++i) out[32 * c + i] = f[i];
}

__global__ __launch_bounds__(kBlock) void tiny_predict_kernel(const double* __restrict__ w,
                                                               const double* __restrict__ feats, int64_t n,
                                                               double* __restrict__ out) {
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= n) return;
  double f[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) f[i] = feats[32 * c + i];
  out[c] = tiny_predict(w, f);
}

// ---- STPF priority trunk (FP32) ------------------------------------------
// Packed weights (floats): w1[128][32] b1[128] w2[128][128] b2[128] wp[128] bp.
constexpr int kH = 128, kF = 32;
constexpr int kSW1 = 0, kSB1 = kSW1 + kH * kF, kSW2 = kSB1 + kH, kSB2 = kSW2 + kH * kH, kSWP = kSB2 + kH,
              kSBP = kSWP + kH, kStpfWeights = kSBP + 1;
static_assert(kStpfWeights == P2C_STPF_WEIGHTS, "STPF weight layout");

__device__ inline float gelu(float v) { return 0.5f * v * (1.0f + erff(v * 0.70710678118654752f)); }

typedef float floatx16 __attribute__((ext_vector_type(16)));

// One wave per 32 rows on FP32 matrix cores (v_mfma_f32_32x32x2f32), both
// layers computed transposed (hidden x rows) so that no lane exchange is
// needed between them.  In a 32x32 accumulator, lane l holds column
// j = l % 32 (a row of the batch) and, in register v, hidden unit
// i = 8 (v / 4) + 4 (l / 32) + v % 4.  Layer 1 takes A = W1 (lane: hidden
// l % 32, feature k) and B = x^T (lane: feature k, row l % 32).  Layer 2
// consumes the layer-1 registers as they are: step (t, v) multiplies the pair
// of hidden units the two lane halves hold in register v of tile t, and the
// A operand gathers the matching W2 columns, so the K order is a permutation
// of 0..127 (FP32 sums in another order than torch's).  The head dot product
// runs on each lane's 64 layer-2 values; the two halves meet by a lane swap.
__global__ __launch_bounds__(256, 2) void stpf_mfma_kernel(const float* __restrict__ w, const float* __restrict__ x,
                                                        int64_t n, float* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t r0 = ((int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6) * 32;
  if (r0 >= n) return;  // whole wave
  const int col = lane & 31;
  const int h = lane >> 5;
  const int64_t row = r0 + col;
  const bool live = row < n;

  floatx16 acc1[4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc1[t][v] = w[kSB1 + 32 * t + 8 * (v >> 2) + 4 * h + (v & 3)];
#pragma unroll
  for (int s = 0; s < kF / 2; ++s) {
    const int k = 2 * s + h;
    const float bx = live ? x[row * kF + k] : 0.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const float aw = w[kSW1 + (32 * t + col) * kF + k];
      acc1[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(aw, bx, acc1[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc1[t][v] = gelu(acc1[t][v]);

  float logit = 0.0f;
#pragma unroll 1
  for (int mo = 0; mo < 4; ++mo) {
    floatx16 acc2;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc2[v] = w[kSB2 + 32 * mo + 8 * (v >> 2) + 4 * h + (v & 3)];
    const float* w2 = w + kSW2 + (32 * mo + col) * kH + 4 * h;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const float4 a4 = *reinterpret_cast<const float4*>(w2 + 32 * t + 8 * g);
        acc2 = __builtin_amdgcn_mfma_f32_32x32x2f32(a4.x, acc1[t][4 * g + 0], acc2, 0, 0, 0);
        acc2 = __builtin_amdgcn_mfma_f32_32x32x2f32(a4.y, acc1[t][4 * g + 1], acc2, 0, 0, 0);
        acc2 = __builtin_amdgcn_mfma_f32_32x32x2f32(a4.z, acc1[t][4 * g + 2], acc2, 0, 0, 0);
        acc2 = __builtin_amdgcn_mfma_f32_32x32x2f32(a4.w, acc1[t][4 * g + 3], acc2, 0, 0, 0);
      }
#pragma unroll
    for (int v = 0; v < 16; ++v)
      logit = fmaf(w[kSWP + 32 * mo + 8 * (v >> 2) + 4 * h + (v & 3)], gelu(acc2[v]), logit);
  }
  logit += __shfl_xor(logit, 32);
  if (h == 0 && live) out[row] = 1.0f / (1.0f + expf(-(logit + w[kSBP])));
}

}  // namespace
}  // namespace p2c

extern "C" {

int p2c_tiny_stpf_score(int family, const double* x0, const double* x1, const int32_t* topo, const int32_t* pairs,
                        int64_t n, const uint64_t* totals, const int32_t* object_a, const int32_t* object_b,
                        int policy, const double* weights, double* scores, uint64_t* group_keys, void* stream) {
  if (family != P2C_FAMILY_POINT_TRIANGLE && family != P2C_FAMILY_EDGE_EDGE)
    return p2c::set_error(P2C_EINVAL, "family must be point-triangle (1) or edge-edge (2)");
  if (policy < P2C_SCORE_LEARNED || policy > P2C_SCORE_RANDOM)
    return p2c::set_error(P2C_EINVAL, "unknown score policy");
  if ((policy == P2C_SCORE_LEARNED || policy == P2C_SCORE_LEARNED_RESIDUAL) && weights == nullptr)
    return p2c::set_error(P2C_EINVAL, "learned policy requested without STPF weights");
  if (n <= 0) return P2C_OK;
  hipLaunchKernelGGL(p2c::tiny_score_kernel, dim3(p2c::grid_for(n, p2c::kBlock)), dim3(p2c::kBlock), 0,
                     p2c::as_stream(stream), family == P2C_FAMILY_POINT_TRIANGLE ? 1 : 0, x0, x1, topo, pairs, n,
                     totals, object_a, object_b, policy, weights, scores, group_keys);
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int p2c_tiny_stpf_features(int family, const double* rows, const uint64_t* index, int64_t n, double* out,
                           void* stream) {
  if (family != P2C_FAMILY_POINT_TRIANGLE && family != P2C_FAMILY_EDGE_EDGE)
    return p2c::set_error(P2C_EINVAL, "family must be point-triangle (1) or edge-edge (2)");
  if (n <= 0) return P2C_OK;
  hipLaunchKernelGGL(p2c::tiny_features_kernel, dim3(p2c::grid_for(n, p2c::kBlock)), dim3(p2c::kBlock), 0,
                     p2c::as_stream(stream), family == P2C_FAMILY_POINT_TRIANGLE ? 1 : 0, rows, index, n, out);
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int p2c_tiny_stpf_predict(const double* weights, const double* features, int64_t n, double* out, void* stream) {
  if (weights == nullptr) return p2c::set_error(P2C_EINVAL, "learned policy requested without STPF weights");
  if (n <= 0) return P2C_OK;
  hipLaunchKernelGGL(p2c::tiny_predict_kernel, dim3(p2c::grid_for(n, p2c::kBlock)), dim3(p2c::kBlock), 0,
                     p2c::as_stream(stream), weights, features, n, out);
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int p2c_stpf_priority(const float* weights, const float* features, int64_t n, float* out, void* stream) {
  if (weights == nullptr) return p2c::set_error(P2C_EINVAL, "STPF weights are required");
  if (n <= 0) return P2C_OK;
  // 32 rows per wave, 4 waves per workgroup
  hipLaunchKernelGGL(p2c::stpf_mfma_kernel, dim3(p2c::grid_for((n + 31) / 32, 4)), dim3(256), 0,
                     p2c::as_stream(stream), weights, features, n, out);
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

}  // extern "C"
