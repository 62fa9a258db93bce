// Reference-parity interval certificate on gfx950: one lane per primitive,
// stackless depth-first walk (cert_ref_math.h), FP64 without contraction.
//
// Replaces the reference's CUDA batch (src/cpp/certificate/certificate_cuda.cu:
// 286-392, 456-552) but follows the CPU engine's semantics
// (certificate_engine.cpp:300-386): the reference CUDA twin stops at the first
// undecided leaf, which can change a collision into "undecided"; this kernel
// keeps walking as the CPU engine does, so its answers equal the CPU's.
#include <cmath>
#include <vector>

#include "cert_ref_math.h"
#include "runtime.h"

namespace p2c {
namespace {

constexpr int kRefBlock = 256;

template <int kFamily>
__global__ __launch_bounds__(kRefBlock) void cert_ref_kernel(
    const p2c_prim* __restrict__ prims, int64_t n, double t0, double t1, double eps_time,
    double eps_space, uint32_t max_depth, p2c_prim_result* __restrict__ out,
    uint32_t* __restrict__ invalid_count) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const p2c_prim p = prims[i];
    p2c_prim_result r;
    if (!prim_valid(p)) {
      init_result(&r, P2C_STATUS_INVALID, kFamily, t0, t1);
      r.status = P2C_STATUS_INVALID;
      r.reason_code = 3;             // kCertificateReasonInvalidInput
      r.next_refinement_mode = 2;    // kRequestGeometry
      if (invalid_count) atomicAdd(invalid_count, 1u);
    } else {
      const RefPrim q = load_prim(p);
      ref_certify<kFamily>(q, p.id_a, p.id_b, t0, t1, eps_time, eps_space, max_depth, &r);
    }
    out[i] = r;
  }
}

template <int kFamily>
__global__ __launch_bounds__(kRefBlock) void cert_ref_distance_kernel(
    const p2c_prim* __restrict__ prims, const double* __restrict__ t, int64_t n,
    double* __restrict__ out) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const RefPrim q = load_prim(prims[i]);
    out[i] = ref_distance<kFamily>(q, t[i]);
  }
}

// ValidateConfig / ValidateInterval (certificate_engine.cpp:224-242), plus the
// depth bound this stackless walk supports.
int check_args(int family, double t0, double t1, const p2c_cert_config* cfg, uint32_t* eff_depth) {
  if (family != P2C_FAMILY_POINT_TRIANGLE && family != P2C_FAMILY_EDGE_EDGE)
    return set_error(P2C_EINVAL, "feature family must be point-triangle (1) or edge-edge (2)");
  if (cfg == nullptr) return set_error(P2C_EINVAL, "CertificateEngineConfig is required");
  if (!std::isfinite(cfg->eps_time) || cfg->eps_time <= 0.0)
    return set_error(P2C_EINVAL, "CertificateEngineConfig.eps_time must be finite and positive");
  if (!std::isfinite(cfg->eps_space) || cfg->eps_space <= 0.0)
    return set_error(P2C_EINVAL, "CertificateEngineConfig.eps_space must be finite and positive");
  if (!std::isfinite(t0) || !std::isfinite(t1))
    return set_error(P2C_EINVAL, "certificate interval endpoints must be finite");
  if (t0 < 0.0 || t1 > 1.0 || t0 > t1)
    return set_error(P2C_EINVAL, "certificate interval must satisfy 0 <= t0 <= t1 <= 1");
  // Depth at which the nominal width (t1 - t0) / 2^d first reaches eps_time;
  // two levels of slack cover midpoint rounding.
  uint32_t d = 0;
  double w = t1 - t0;
  while (w > cfg->eps_time && d < 80) {
    w *= 0.5;
    ++d;
  }
  uint32_t eff = cfg->max_subdivision_depth;
  if (d + 2 < eff) eff = d + 2;
  if (eff > 63)
    return set_error(P2C_EUNSUPPORTED,
                     "subdivision deeper than 63 levels is not supported by the HIP certificate");
  *eff_depth = cfg->max_subdivision_depth;
  return P2C_OK;
}

const char* invalid_message(int family, const p2c_prim& p) {
  const bool pt = family == P2C_FAMILY_POINT_TRIANGLE;
  if (p.id_a < 0 || p.id_b < 0)
    return pt ? "point-triangle primitive ids are required" : "edge-edge primitive ids are required";
  return pt ? "point-triangle primitive trajectories must be finite"
            : "edge-edge primitive trajectories must be finite";
}

}  // namespace
}  // namespace p2c

extern "C" {

int p2c_cert_ref_batch(int family, const p2c_prim* prims, int64_t n, double t0, double t1,
                       const p2c_cert_config* cfg, p2c_prim_result* out,
                       uint32_t* invalid_count, void* stream) {
  uint32_t depth = 0;
  if (int rc = p2c::check_args(family, t0, t1, cfg, &depth)) return rc;
  if (n < 0) return p2c::set_error(P2C_EINVAL, "primitive count must be non-negative");
  if (n == 0) return P2C_OK;
  if (prims == nullptr || out == nullptr)
    return p2c::set_error(P2C_EINVAL, "primitive and result buffers are required");
  const unsigned grid = p2c::grid_for(n, p2c::kRefBlock);
  hipStream_t s = p2c::as_stream(stream);
  if (family == P2C_FAMILY_POINT_TRIANGLE) {
    hipLaunchKernelGGL(p2c::cert_ref_kernel<P2C_FAMILY_POINT_TRIANGLE>, dim3(grid),
                       dim3(p2c::kRefBlock), 0, s, prims, n, t0, t1, cfg->eps_time,
                       cfg->eps_space, depth, out, invalid_count);
  } else {
    hipLaunchKernelGGL(p2c::cert_ref_kernel<P2C_FAMILY_EDGE_EDGE>, dim3(grid),
                       dim3(p2c::kRefBlock), 0, s, prims, n, t0, t1, cfg->eps_time,
                       cfg->eps_space, depth, out, invalid_count);
  }
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int p2c_cert_ref_distance(int family, const p2c_prim* prims, const double* t, int64_t n,
                          double* out, void* stream) {
  if (family != P2C_FAMILY_POINT_TRIANGLE && family != P2C_FAMILY_EDGE_EDGE)
    return p2c::set_error(P2C_EINVAL, "feature family must be point-triangle (1) or edge-edge (2)");
  if (n <= 0) return P2C_OK;
  const unsigned grid = p2c::grid_for(n, p2c::kRefBlock);
  hipStream_t s = p2c::as_stream(stream);
  if (family == P2C_FAMILY_POINT_TRIANGLE)
    hipLaunchKernelGGL(p2c::cert_ref_distance_kernel<P2C_FAMILY_POINT_TRIANGLE>, dim3(grid),
                       dim3(p2c::kRefBlock), 0, s, prims, t, n, out);
  else
    hipLaunchKernelGGL(p2c::cert_ref_distance_kernel<P2C_FAMILY_EDGE_EDGE>, dim3(grid),
                       dim3(p2c::kRefBlock), 0, s, prims, t, n, out);
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int p2c_cert_ref_batch_host(int family, const p2c_prim* prims, int64_t n, double t0, double t1,
                            const p2c_cert_config* cfg, p2c_prim_result* out) {
  uint32_t depth = 0;
  if (int rc = p2c::check_args(family, t0, t1, cfg, &depth)) return rc;
  if (n < 0) return p2c::set_error(P2C_EINVAL, "primitive count must be non-negative");
  if (n == 0) return P2C_OK;
  // Host-side validation first: the reference's CPU path rejects the batch with
  // the first failing primitive's message.
  for (int64_t i = 0; i < n; ++i)
    if (!p2c::prim_valid(prims[i]))
      return p2c::set_error(P2C_EINVAL, p2c::invalid_message(family, prims[i]));
  p2c_prim* d_prims = nullptr;
  p2c_prim_result* d_out = nullptr;
  const size_t in_bytes = sizeof(p2c_prim) * size_t(n);
  const size_t out_bytes = sizeof(p2c_prim_result) * size_t(n);
  hipError_t e = hipMalloc(&d_prims, in_bytes);
  if (e != hipSuccess) return p2c::hip_fail(e, "hipMalloc primitives");
  e = hipMalloc(&d_out, out_bytes);
  if (e != hipSuccess) {
    (void)hipFree(d_prims);
    return p2c::hip_fail(e, "hipMalloc results");
  }
  int rc = P2C_OK;
  e = hipMemcpy(d_prims, prims, in_bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) rc = p2c::hip_fail(e, "hipMemcpy primitives H2D");
  if (rc == P2C_OK) rc = p2c_cert_ref_batch(family, d_prims, n, t0, t1, cfg, d_out, nullptr, nullptr);
  if (rc == P2C_OK) {
    e = hipMemcpy(out, d_out, out_bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = p2c::hip_fail(e, "hipMemcpy results D2H");
  }
  (void)hipFree(d_prims);
  (void)hipFree(d_out);
  return rc;
}

int p2c_cert_ref_eval_cpu(int family, const p2c_prim* prim, double t0, double t1,
                          const p2c_cert_config* cfg, p2c_prim_result* out) {
  uint32_t depth = 0;
  if (int rc = p2c::check_args(family, t0, t1, cfg, &depth)) return rc;
  if (prim == nullptr || out == nullptr)
    return p2c::set_error(P2C_EINVAL, "primitive and result pointers are required");
  if (!p2c::prim_valid(*prim)) return p2c::set_error(P2C_EINVAL, p2c::invalid_message(family, *prim));
  const p2c::RefPrim q = p2c::load_prim(*prim);
  if (family == P2C_FAMILY_POINT_TRIANGLE)
    p2c::ref_certify<P2C_FAMILY_POINT_TRIANGLE>(q, prim->id_a, prim->id_b, t0, t1, cfg->eps_time,
                                                cfg->eps_space, depth, out);
  else
    p2c::ref_certify<P2C_FAMILY_EDGE_EDGE>(q, prim->id_a, prim->id_b, t0, t1, cfg->eps_time,
                                           cfg->eps_space, depth, out);
  return P2C_OK;
}

}  // extern "C"
