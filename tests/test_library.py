"""CPU-side checks of the shipped library: it loads, exports every symbol the
C header declares, its host certificate path equals the oracle bit-for-bit,
and the p2cccd_cpp surface keeps the reference's names, messages and host
rules (src/tests/cpp/test_certificate_engine.cpp:180-337,
src/tests/python/test_cpp_bindings.py:20-215).  No kernel is launched."""
import re

import numpy as np
import pytest

import oracle
from p2cccd import _native, workloads

HEADER = oracle.HERE.parent / "include" / "p2c_hip.h"


def test_library_exports_every_declared_symbol():
    text = HEADER.read_text()
    declared = set(re.findall(r"^\s*(?:int|size_t|const char\*|void)\s+(p2c_\w+)\(", text, re.M))
    assert len(declared) >= 18
    missing = [name for name in sorted(declared) if not hasattr(_native.lib, name)]
    assert missing == []


def test_host_certificate_equals_oracle_bit_exact():
    qs = workloads.make_query_set(800, 800, seed=1234)
    cfg = _native.cert_config()
    for fam, x in ((1, qs.vf), (2, qs.ee)):
        mine = _native.cert_ref_eval_cpu(fam, _native.prims_from_positions(x), 0.0, 1.0, cfg)
        ref = oracle.ref_cert(fam, oracle.prims_from_positions(x))
        for f in ("status", "toi_upper", "safe_margin_lb", "interval_t0", "interval_t1", "reason_code",
                  "witness_id_a", "witness_id_b", "covered_feature_mask", "nodes"):
            assert np.array_equal(mine[f].astype(np.float64), ref[f].astype(np.float64)), (fam, f)


def test_host_certificate_matches_golden():
    g = np.load(oracle.HERE.parent / "tests" / "golden" / "ref_cert_golden.npz")
    for i in range(0, len(g["family"]), 7):
        t0, t1, et, es, d = g["params"][i]
        p = _native.prims_from_positions(g["x"][i][None], ids=np.array([[2 * i + 10, 2 * i + 11]]))
        r = _native.cert_ref_eval_cpu(int(g["family"][i]), p, t0, t1, _native.cert_config(et, es, int(d)))[0]
        assert r["status"] == g["status"][i]
        assert r["toi_upper"] == g["toi_upper"][i] and r["safe_margin_lb"] == g["safe_margin_lb"][i]
        assert (r["interval_t0"], r["interval_t1"]) == (g["interval_t0"][i], g["interval_t1"][i])


def test_c_abi_rejects_bad_config_with_reference_text():
    p = _native.prims_from_positions(np.zeros((1, 8, 3)))
    with pytest.raises(_native.NativeError, match="eps_time must be finite and positive"):
        _native.cert_ref_eval_cpu(1, p, 0.0, 1.0, _native.cert_config(eps_time=0.0))
    with pytest.raises(_native.NativeError, match="0 <= t0 <= t1 <= 1"):
        _native.cert_ref_eval_cpu(1, p, 0.5, 0.25, _native.cert_config())
    bad = p.copy()
    bad["id_a"] = -1
    with pytest.raises(_native.NativeError, match="point-triangle primitive ids are required"):
        _native.cert_ref_eval_cpu(1, bad, 0.0, 1.0, _native.cert_config())
    bad = p.copy()
    bad["v"][0, 2]["position_t1"][1] = np.nan
    with pytest.raises(_native.NativeError, match="edge-edge primitive trajectories must be finite"):
        _native.cert_ref_eval_cpu(2, bad, 0.0, 1.0, _native.cert_config())


def _sliding_graze(h_ulps: int, u: float) -> np.ndarray:
    """A point descending onto the plane of a triangle that slides under it,
    ending h_ulps doubles above eps_space = 1e-6 at t = 1: the motion radius
    keeps the last interval [t, 1] undecided down to one double spacing, where
    eps_time 1e-17 can no longer end the walk and only the depth cap does."""
    h = 1e-6
    for _ in range(h_ulps):
        h = np.nextafter(h, 1.0)
    x = np.zeros((1, 8, 3))
    tri = np.array([(0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)])
    x[0, 0], x[0, 4] = (0.25, 0.25, h + 0.1), (0.25, 0.25, h)
    x[0, 1:4], x[0, 5:8] = tri, tri - (u, 0.0, 0.0)
    return x


def test_walk_depth_guard_refuses_walks_past_63_levels():
    # eps_time below the spacing of doubles near t = 1: an interval of one
    # spacing splits into a copy of itself, so max_subdivision_depth alone
    # bounds the walk; past 63 levels the path word would overflow.
    x = _sliding_graze(1, 0.1)
    p = _native.prims_from_positions(x)
    for depth in (64, 65, 200):
        cfg = _native.cert_config(1e-17, 1e-6, depth)
        with pytest.raises(_native.NativeError, match="deeper than 63 levels") as e:
            _native.cert_ref_eval_cpu(1, p, 0.0, 1.0, cfg)
        assert e.value.code == -4
        out = np.zeros(1, _native.PRIM_RESULT_DTYPE)
        with pytest.raises(_native.NativeError, match="deeper than 63 levels"):  # refused before any device call
            _native.call("p2c_cert_ref_batch_host", 1, _native.ptr(p), 1, 0.0, 1.0, _native.ptr(cfg),
                         _native.ptr(out))
    # a large cap with eps_time well above the spacing: widths end the walk
    # long before 63 levels, so it is accepted and agrees with the oracle
    qs = workloads.make_query_set(60, 60, seed=5)
    for fam, xs in ((1, qs.vf), (2, qs.ee)):
        mine = _native.cert_ref_eval_cpu(fam, _native.prims_from_positions(xs), 0.0, 1.0,
                                         _native.cert_config(1e-6, 1e-6, 500))
        ref = oracle.ref_cert(fam, oracle.prims_from_positions(xs), eps_time=1e-6, max_depth=500)
        assert np.array_equal(mine["status"], ref["status"]) and np.array_equal(mine["nodes"], ref["nodes"])


@pytest.mark.parametrize("u,nodes", [(0.1, 127), (0.3, 231)])
def test_walk_reaches_the_63_level_cap_like_the_reference(u, nodes):
    # 63 levels is the deepest walk allowed: the copy chain ends there,
    # UNDECIDED, as the reference engine ends it (2 * 63 + 1 nodes for u 0.1).
    x = _sliding_graze(1, u)
    op = oracle.prims_from_positions(x)
    for flags, ref in ((0, oracle.ref_cert(1, op, eps_time=1e-17, max_depth=63)),
                       (1, oracle.ref_cert_twin(1, op, eps_time=1e-17, max_depth=63))):
        cfg = _native.cert_config(1e-17, 1e-6, 63)
        cfg["flags"] = flags
        mine = _native.cert_ref_eval_cpu(1, _native.prims_from_positions(x), 0.0, 1.0, cfg)
        assert mine["status"][0] == 2 and mine["nodes"][0] == ref["nodes"][0]
        for f in ("interval_t0", "interval_t1", "toi_upper", "reason_code"):
            assert mine[f][0] == ref[f][0], (flags, f)
    assert oracle.ref_cert(1, op, eps_time=1e-17, max_depth=63)["nodes"][0] == nodes
    if oracle.ref_available():
        r = oracle.ref_cert_compiled(1, op, eps_time=1e-17, max_depth=63)
        assert r["status"][0] == 2 and r["toi_upper"][0] == 1.0


def test_host_certificate_follows_every_walk_flag():
    # flags select the walk on the host entry as on the kernels: the CUDA
    # twin's (UNDECIDED on the first unsplittable interval) and the Python
    # twin's, each against its oracle restatement.
    qs = workloads.make_query_set(500, 500, seed=2024)
    for fam, x in ((1, qs.vf), (2, qs.ee)):
        p, op = _native.prims_from_positions(x), oracle.prims_from_positions(x)
        mine = _native.cert_ref_eval_cpu(fam, p, 0.0, 1.0, _native.cert_config(cuda_twin=True))
        ref = oracle.ref_cert_twin(fam, op)
        for f in ("status", "toi_upper", "safe_margin_lb", "interval_t0", "interval_t1", "nodes"):
            assert np.array_equal(mine[f].astype(np.float64), ref[f].astype(np.float64)), (fam, f)
        cfg = _native.cert_config(1e-5, 1e-8, 32)
        cfg["flags"] = 2
        mine = _native.cert_ref_eval_cpu(fam, p, 0.0, 1.0, cfg)
        ref = oracle.ref_cert_py(fam, op)
        for f in ("status", "toi_upper", "safe_margin_lb", "nodes"):
            assert np.array_equal(mine[f].astype(np.float64), ref[f].astype(np.float64)), (fam, f)
        assert np.array_equal(mine["pad"][:, 0].astype(np.int64), ref["reason_code"].astype(np.int64))
    p = _native.prims_from_positions(qs.vf[:1])
    for flags, msg in ((3, "choose one certificate walk"), (4, "unknown certificate flags"),
                       (0x80000000, "unknown certificate flags")):
        cfg = _native.cert_config()
        cfg["flags"] = flags
        with pytest.raises(_native.NativeError, match=msg):
            _native.cert_ref_eval_cpu(1, p, 0.0, 1.0, cfg)


# ---- p2cccd_cpp surface (host rules only) ---------------------------------
cpp = pytest.importorskip("p2cccd_cpp")


def _item(m, mask, wid=9001):
    it = m.ExactWorkItem()
    it.work_item_id, it.parent_candidate_id, it.query_id = wid, 77, 1234
    it.patch_a_id, it.patch_b_id = 1, 2
    it.interval_t0, it.interval_t1 = 0.0, 1.0
    it.feature_family_mask = mask
    it.priority_score = 1.0
    it.source = m.ProposalSource.RAW
    return it


def test_module_surface_matches_reference_names():
    for name in ("is_cuda_exact_built", "cuda_binding_status", "evaluate_point_triangle_batch_cuda",
                 "evaluate_edge_edge_batch_cuda", "cross_check_cpu_cuda_exact", "evaluate_point_triangle_interval",
                 "evaluate_edge_edge_interval", "evaluate_certificate_query_cpu", "process_exact_work_queue_cpu",
                 "validate_exact_work_queue_coverage", "generate_conservative_refinement_work_items",
                 "LinearVertexTrajectory", "PointTriangleIntervalPrimitive", "EdgeEdgeIntervalPrimitive",
                 "CertificateEngineConfig", "PrimitiveIntervalResult", "CertificateEngine"):
        assert hasattr(cpp, name), name
    st = cpp.cuda_binding_status()
    assert {"cuda_exact_built", "host_batch_exact_api", "device_pointer_abi", "backend_name",
            "safety_policy"} <= set(st)
    assert cpp.FEATURE_FAMILY_POINT_TRIANGLE == 1 and cpp.FEATURE_FAMILY_EDGE_EDGE == 2
    assert cpp.CERTIFICATE_REASON_MAX_SUBDIVISION_DEPTH == 2 and cpp.EXACT_AUDIT_DEQUEUED == 1


def test_interval_kat_through_module_host_path():
    # test_certificate_engine.cpp:97-115 through p2cccd_cpp (host CPU entry)
    m = cpp
    p = m.PointTriangleIntervalPrimitive()
    p.point_id, p.triangle_id = 10, 20

    def v(i, a, b):
        t = m.LinearVertexTrajectory()
        t.feature_id, t.position_t0, t.position_t1 = i, a, b
        return t

    p.point = v(10, [0.25, 0.25, 1.0], [0.25, 0.25, -3.0])
    p.triangle_v0 = v(101, [0, 0, 0], [0, 0, 0])
    p.triangle_v1 = v(102, [1, 0, 0], [1, 0, 0])
    p.triangle_v2 = v(103, [0, 1, 0], [0, 1, 0])
    cfg = m.CertificateEngineConfig()
    cfg.eps_time, cfg.eps_space, cfg.max_subdivision_depth = 1e-5, 1e-6, 32
    r = m.evaluate_point_triangle_interval(p, 0.0, 1.0, cfg)
    assert r.status == m.CertificateStatus.COLLISION and abs(r.toi_upper - 0.25) <= 1e-12
    assert (r.witness_id_a, r.witness_id_b) == (10, 20)
    cfg.eps_time = -1.0
    with pytest.raises(ValueError, match="eps_time must be finite and positive"):
        m.evaluate_point_triangle_interval(p, 0.0, 1.0, cfg)


def test_refinement_children_reference_kat():
    # test_certificate_engine.cpp:282-337
    m = cpp
    parent = _item(m, 1)
    col = m.CertificateResult()
    col.work_item_id, col.query_id = parent.work_item_id, parent.query_id
    col.status = m.CertificateStatus.COLLISION
    col.interval_t0, col.interval_t1, col.toi_upper = 0.0, 1.0, 0.25
    col.witness_family, col.witness_id_a, col.witness_id_b = 1, 10, 20
    cfg = m.ExactRefinementConfig()
    cfg.first_child_work_item_id, cfg.min_interval_width = 10001, 1e-5
    ch = m.generate_conservative_refinement_work_items(parent, col, cfg)
    assert len(ch) == 1 and ch[0].interval_t0 == 0.0 and ch[0].interval_t1 == 0.25
    assert ch[0].source == m.ProposalSource.REFINED and ch[0].work_item_id == 10001
    und = m.CertificateResult()
    und.work_item_id, und.query_id = parent.work_item_id, parent.query_id
    und.status = m.CertificateStatus.UNDECIDED
    und.interval_t0, und.interval_t1, und.toi_upper = 0.0, 1.0, 1.0
    und.reason_code = m.CERTIFICATE_REASON_MAX_SUBDIVISION_DEPTH
    und.next_refinement_mode = m.CertificateRefinementMode.BISECT_INTERVAL
    ch = m.generate_conservative_refinement_work_items(parent, und, cfg)
    assert len(ch) == 2 and ch[0].interval_t1 == 0.5 and ch[1].interval_t0 == 0.5
    assert (ch[0].work_item_id, ch[1].work_item_id) == (10001, 10002)


def test_coverage_guard_and_validators():
    m = cpp
    q = m.ExactCertificateQuery()
    q.work_item = _item(m, 1)
    incomplete = m.ExactWorkQueueResult()
    incomplete.processed_count = 1
    with pytest.raises(ValueError, match="exactly one certificate"):
        m.validate_exact_work_queue_coverage([q], incomplete)
    bad = _item(m, 0)
    with pytest.raises(ValueError, match="feature_family_mask is required"):
        m.validate_exact_work_item(bad)
    r = m.CertificateResult()
    r.work_item_id, r.query_id = 1, 1
    r.status = m.CertificateStatus.SEPARATION
    with pytest.raises(ValueError, match="covered_feature_mask is required for separation"):
        m.validate_certificate_result(r)
    leg = m.CertificateEngine().evaluate_work_item(_item(m, 1))
    assert leg.status == m.CertificateStatus.UNDECIDED
    assert leg.next_refinement_mode == m.CertificateRefinementMode.REQUEST_GEOMETRY
