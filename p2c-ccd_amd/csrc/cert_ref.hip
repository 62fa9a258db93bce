// Reference-parity interval certificate on gfx950, FP64 without contraction:
// the per-batch entry points over the walk kernels of cert_ref_walk.h.
//
// Replaces the reference's CUDA batch (src/cpp/certificate/certificate_cuda.cu:
// 286-392, 456-552).  The walk is a compile-time mode chosen by
// p2c_cert_config.flags: the CPU engine's (certificate_engine.cpp:300-386,
// the default: an interval that may not be split is noted and the search goes
// on), the CUDA twin's (certificate_cuda.cu:286-346: UNDECIDED on the first
// such interval) or the Python twin's (pure_exact_cpu.py:98-244).  The same
// walk functions run on the host for p2c_cert_ref_eval_cpu.
#include <cmath>
#include <vector>

#include "cert_ref_walk.h"
#include "runtime.h"

namespace p2c {
namespace {

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

// ValidateConfig / ValidateInterval (certificate_engine.cpp:224-242), the
// walk flags, and the depth bound of this walk: a node's path is one 64-bit
// word, so no node may lie deeper than 63 levels.  An interval stops
// splitting at max_subdivision_depth or once its width b - a is at most
// eps_time.  With u the spacing of doubles just above max(t0, t1), a split of
// [a, b] leaves children no wider than (b - a) / 2 + u / 2 (the rounded
// midpoint), so the widths after k splits are at most (t1 - t0) / 2^k + u and
// the computed width at most that plus u / 2; when eps_time >= 4u every
// interval is therefore done by depth d + 1, d the first depth at which the
// nominal width (t1 - t0) / 2^d reaches eps_time.  Below 4u an interval can
// stop shrinking (a midpoint that rounds to an end reproduces its parent) and
// only max_subdivision_depth ends the walk.  Either way the deepest reachable
// node must be at most 63 deep, or the call is refused; *walk_depth is then
// min(max_subdivision_depth, 63), which changes no answer.
int check_args(int family, double t0, double t1, const p2c_cert_config* cfg, uint32_t* walk_depth,
               int* mode) {
  if (family != P2C_FAMILY_POINT_TRIANGLE && family != P2C_FAMILY_EDGE_EDGE)
    return set_error(P2C_EINVAL, "feature family must be point-triangle (1) or edge-edge (2)");
  if (cfg == nullptr) return set_error(P2C_EINVAL, "CertificateEngineConfig is required");
  if (cfg->flags & ~(P2C_CERT_CUDA_TWIN | P2C_CERT_PYTHON_TWIN))
    return set_error(P2C_EINVAL, "unknown certificate flags");
  if ((cfg->flags & P2C_CERT_CUDA_TWIN) && (cfg->flags & P2C_CERT_PYTHON_TWIN))
    return set_error(P2C_EINVAL, "choose one certificate walk: CUDA twin or Python twin");
  if (!std::isfinite(cfg->eps_time) || cfg->eps_time <= 0.0)
    return set_error(P2C_EINVAL, "CertificateEngineConfig.eps_time must be finite and positive");
  if (!std::isfinite(cfg->eps_space) || cfg->eps_space <= 0.0)
    return set_error(P2C_EINVAL, "CertificateEngineConfig.eps_space must be finite and positive");
  if (!std::isfinite(t0) || !std::isfinite(t1))
    return set_error(P2C_EINVAL, "certificate interval endpoints must be finite");
  if (t0 < 0.0 || t1 > 1.0 || t0 > t1)
    return set_error(P2C_EINVAL, "certificate interval must satisfy 0 <= t0 <= t1 <= 1");
  const double top = std::fmax(std::fabs(t0), std::fabs(t1));
  const double spacing = std::nextafter(top, HUGE_VAL) - top;
  uint32_t reach = cfg->max_subdivision_depth;
  if (cfg->eps_time >= 4.0 * spacing) {
    uint32_t d = 0;
    double w = t1 - t0;
    while (w > cfg->eps_time && d < 80) {
      w *= 0.5;
      ++d;
    }
    if (d + 2 < reach) reach = d + 2;
  }
  if (reach > 63)
    return set_error(P2C_EUNSUPPORTED,
                     "subdivision deeper than 63 levels is not supported by the HIP certificate");
  *walk_depth = cfg->max_subdivision_depth < 63u ? cfg->max_subdivision_depth : 63u;
  *mode = (cfg->flags & P2C_CERT_CUDA_TWIN)     ? kModeCudaTwin
          : (cfg->flags & P2C_CERT_PYTHON_TWIN) ? kModePython
                                                : kModeEngine;
  return P2C_OK;
}

// One primitive's certificate on the host: the light pass's walk with no
// node budget, so the same functions and operation order as the kernels.
template <int kFamily, int kMode>
void host_certify(const RefPrim& q, const p2c_prim& p, const RefCfg& c, p2c_prim_result* out) {
  double vn[4];
  vertex_norms(q, vn);
  const double d0 = ref_distance<kFamily>(q, c.t0, kMode == kModePython);
  const double d1 = ref_distance<kFamily>(q, c.t1, kMode == kModePython);
  const RefWalk w = walk_subtree<kFamily, kMode>(q, vn, c, 0, 0, d0, d1, 0xFFFFFFFFu);
  write_result<kFamily>(c, p.id_a, p.id_b, w.state == kWalkCollision, w.limited, w.a, w.b, w.toi, w.undecided,
                        w.min_margin, w.nodes, w.max_depth, out);
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
  int mode = 0;
  if (int rc = p2c::check_args(family, t0, t1, cfg, &depth, &mode)) return rc;
  if (n < 0) return p2c::set_error(P2C_EINVAL, "primitive count must be non-negative");
  if (n == 0) return P2C_OK;
  if (prims == nullptr || out == nullptr)
    return p2c::set_error(P2C_EINVAL, "primitive and result buffers are required");
  const p2c::UniformSrc src{prims, p2c::RefCfg{t0, t1, cfg->eps_time, cfg->eps_space, depth}};
  p2c::launch_walk(family, mode, src, n, out, invalid_count, p2c::as_stream(stream));
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
  int mode = 0;
  if (int rc = p2c::check_args(family, t0, t1, cfg, &depth, &mode)) return rc;
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
  int mode = 0;
  if (int rc = p2c::check_args(family, t0, t1, cfg, &depth, &mode)) return rc;
  if (prim == nullptr || out == nullptr)
    return p2c::set_error(P2C_EINVAL, "primitive and result pointers are required");
  if (!p2c::prim_valid(*prim)) return p2c::set_error(P2C_EINVAL, p2c::invalid_message(family, *prim));
  const p2c::RefCfg c{t0, t1, cfg->eps_time, cfg->eps_space, depth};
  const p2c::RefPrim q = p2c::load_prim(*prim);
  if (family == P2C_FAMILY_POINT_TRIANGLE) {
    if (mode == p2c::kModeEngine) p2c::host_certify<P2C_FAMILY_POINT_TRIANGLE, p2c::kModeEngine>(q, *prim, c, out);
    else if (mode == p2c::kModeCudaTwin)
      p2c::host_certify<P2C_FAMILY_POINT_TRIANGLE, p2c::kModeCudaTwin>(q, *prim, c, out);
    else p2c::host_certify<P2C_FAMILY_POINT_TRIANGLE, p2c::kModePython>(q, *prim, c, out);
  } else {
    if (mode == p2c::kModeEngine) p2c::host_certify<P2C_FAMILY_EDGE_EDGE, p2c::kModeEngine>(q, *prim, c, out);
    else if (mode == p2c::kModeCudaTwin)
      p2c::host_certify<P2C_FAMILY_EDGE_EDGE, p2c::kModeCudaTwin>(q, *prim, c, out);
    else p2c::host_certify<P2C_FAMILY_EDGE_EDGE, p2c::kModePython>(q, *prim, c, out);
  }
  return P2C_OK;
}

}  // extern "C"
