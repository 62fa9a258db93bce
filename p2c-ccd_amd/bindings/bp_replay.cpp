// p2cccd_cpp: a host replay of the broad phase's tree walks, for the
// broad-phase roofline.  Given the trees, query boxes, query orders, counts
// and resume nodes the device left in its workspace (p2c_bp_workspace_layout),
// it walks every query exactly as the count and emit kernels do (escape-linked
// pre-order: on overlap go to `left` unless a leaf, else to `escape`) and
// counts the nodes each visits.  Every visited node is one 64-byte node read,
// the traversal term of the broad phase's algorithmic bytes.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace py = pybind11;

namespace {

struct HostNode {
  double mn[3];
  double mx[3];
  int32_t left, right, escape, prim;
};
static_assert(sizeof(HostNode) == 64, "node layout");

struct HostBox {
  double mn[3];
  double mx[3];
};

// the device's inclusive FP64 overlap test
inline bool Overlap(const HostBox& q, const HostNode& n) {
  return q.mn[0] <= n.mx[0] && n.mn[0] <= q.mx[0] && q.mn[1] <= n.mx[1] && n.mn[1] <= q.mx[1] &&
         q.mn[2] <= n.mx[2] && n.mn[2] <= q.mx[2];
}

uint64_t Walk(const HostNode* nodes, int32_t start, const HostBox& q) {
  uint64_t visits = 0;
  for (int32_t node = start; node >= 0;) {
    const HostNode& n = nodes[node];
    ++visits;
    node = (Overlap(q, n) && n.left >= 0) ? n.left : n.escape;
  }
  return visits;
}

template <typename T>
const T* Data(const py::array& a, size_t count, const char* what) {
  if (size_t(a.nbytes()) < count * sizeof(T)) throw py::value_error(std::string(what) + " is too small");
  return reinterpret_cast<const T*>(a.data());
}

// {count pass, emit pass} visits of the vertex queries and of the edge queries
py::dict TraversalVisits(py::array fnodes, int64_t nf, py::array enodes, int64_t ne, py::array vbox, py::array ebox,
                         py::array qorder, py::array eorder, py::array vertex_object, py::array counts,
                         py::array resume, int64_t nv, int64_t slots) {
  const auto* F = Data<HostNode>(fnodes, size_t(std::max<int64_t>(2 * nf - 1, 0)), "fnodes");
  const auto* E = Data<HostNode>(enodes, size_t(std::max<int64_t>(2 * ne - 1, 0)), "enodes");
  const auto* VB = Data<HostBox>(vbox, size_t(nv), "vbox");
  const auto* EB = Data<HostBox>(ebox, size_t(ne), "ebox");
  const auto* QO = Data<uint32_t>(qorder, size_t(nv), "qorder");
  const auto* EO = Data<uint32_t>(eorder, size_t(ne), "eorder");
  const auto* VO = Data<int32_t>(vertex_object, size_t(nv), "vertex_object");
  const auto* C = Data<uint32_t>(counts, size_t(nv + ne), "counts");
  const auto* R = Data<int32_t>(resume, size_t(nv + ne), "resume");
  const int64_t total = nv + ne;
  const unsigned threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<uint64_t> acc(size_t(threads) * 4, 0);
  py::gil_scoped_release release;
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < threads; ++t) {
    pool.emplace_back([&, t] {
      uint64_t* a = &acc[size_t(t) * 4];
      for (int64_t s = t; s < total; s += threads) {
        if (s < nv) {
          const uint32_t v = QO[s];
          if (VO[v] < 0 || nf == 0) continue;
          a[0] += Walk(F, 0, VB[v]);
          if (C[s] > uint64_t(slots)) a[1] += Walk(F, slots > 0 ? R[s] : 0, VB[v]);
        } else {
          const int64_t e = s - nv;
          const HostBox& q = EB[EO[e]];
          const int32_t start = E[ne - 1 + e].escape;
          a[2] += Walk(E, start, q);
          if (C[s] > uint64_t(slots)) a[3] += Walk(E, slots > 0 ? R[s] : start, q);
        }
      }
    });
  }
  for (auto& th : pool) th.join();
  uint64_t sum[4] = {0, 0, 0, 0};
  for (unsigned t = 0; t < threads; ++t)
    for (int k = 0; k < 4; ++k) sum[k] += acc[size_t(t) * 4 + k];
  py::gil_scoped_acquire acquire;
  py::dict d;
  d["vf_count_visits"] = sum[0];
  d["vf_emit_visits"] = sum[1];
  d["ee_count_visits"] = sum[2];
  d["ee_emit_visits"] = sum[3];
  return d;
}

}  // namespace

void BindReplay(py::module_& m) {
  m.def("lbvh_traversal_visits", &TraversalVisits, py::arg("fnodes"), py::arg("nf"), py::arg("enodes"),
        py::arg("ne"), py::arg("vbox"), py::arg("ebox"), py::arg("qorder"), py::arg("eorder"),
        py::arg("vertex_object"), py::arg("counts"), py::arg("resume"), py::arg("nv"), py::arg("slots"),
        "Nodes the broad phase's count and emit walks visit, replayed on the host from the trees and query "
        "data the device left in its workspace.");
}
