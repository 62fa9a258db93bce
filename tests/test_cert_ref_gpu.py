"""gfx950 reference-parity certificate against the oracle: bit-exact status,
TOI, margin, interval, witness, reason and node count (FP64, no contraction),
on the 20k-query plain workload, the reference's golden vectors, and through
the p2cccd_cpp host-batch API (src/tests/cpp/test_certificate_cuda.cpp:74-100,
src/tests/cpp/test_certificate_engine.cpp:220-280).  The deep configurations
(trees of thousands of nodes) run through the wave pass, whose node counts
must equal the sequential walk's."""
import numpy as np
import pytest

import oracle
from p2cccd import _native, workloads

pytestmark = pytest.mark.gpu

CMP = ("status", "toi_upper", "safe_margin_lb", "interval_t0", "interval_t1", "reason_code",
       "witness_family", "witness_id_a", "witness_id_b", "covered_feature_mask", "next_refinement_mode", "nodes")


def _device_batch(torch, fam, prims, t0=0.0, t1=1.0, cfg=None):
    cfg = _native.cert_config() if cfg is None else cfg
    d_in = torch.from_numpy(prims.view(np.uint8).copy()).cuda()
    d_out = torch.zeros(prims.shape[0] * _native.PRIM_RESULT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    d_bad = torch.zeros(1, dtype=torch.int32, device="cuda")
    _native.call("p2c_cert_ref_batch", fam, d_in.data_ptr(), prims.shape[0], t0, t1, _native.ptr(cfg),
                 d_out.data_ptr(), d_bad.data_ptr(), _native.stream_handle())
    torch.cuda.synchronize()
    return d_out.cpu().numpy().view(_native.PRIM_RESULT_DTYPE), int(d_bad.item())


def _same(res, ref, fields=CMP):
    for f in fields:
        a, b = res[f].astype(np.float64), ref[f].astype(np.float64)
        bad = np.nonzero(a != b)[0]
        assert bad.size == 0, f"{f}: {bad.size} mismatches, first {bad[:5]}"


def test_fp64_distance_bit_exact(gpu):
    import torch

    qs = workloads.make_query_set(4000, 4000, seed=5)
    rng = np.random.default_rng(3)
    for fam, x in ((1, qs.vf), (2, qs.ee)):
        n = x.shape[0]
        t = rng.uniform(0, 1, n)
        t[: n // 4] = np.round(t[: n // 4] * 1024) / 1024
        p = _native.prims_from_positions(x)
        d_p = torch.from_numpy(p.view(np.uint8).copy()).cuda()
        d_t = torch.from_numpy(t).cuda()
        d_o = torch.empty(n, dtype=torch.float64, device="cuda")
        _native.call("p2c_cert_ref_distance", fam, d_p.data_ptr(), d_t.data_ptr(), n, d_o.data_ptr(),
                     _native.stream_handle())
        got = d_o.cpu().numpy()
        op = oracle.prims_from_positions(x)
        fn = oracle.lib().orc_pt_distance if fam == 1 else oracle.lib().orc_ee_distance
        ref = np.array([fn(op[i:i + 1].ctypes.data, t[i]) for i in range(n)])
        assert np.array_equal(got, ref)


def test_plain_workload_bit_exact_vs_oracle(gpu):
    import torch

    qs = workloads.make_query_set(10_000, 10_000)
    for fam, x in ((1, qs.vf), (2, qs.ee)):
        res, bad = _device_batch(torch, fam, _native.prims_from_positions(x))
        assert bad == 0
        ref = oracle.ref_cert(fam, oracle.prims_from_positions(x))
        _same(res, ref)
        assert set(np.unique(res["status"])) == {0, 1, 2}


@pytest.mark.parametrize("kw", [dict(), dict(t0=0.125, t1=0.875), dict(eps_time=1e-6, eps_space=1e-4, depth=20),
                                dict(depth=3), dict(eps_time=1e-8, eps_space=1e-6, depth=40)])
def test_configs_bit_exact(gpu, kw):
    import torch

    qs = workloads.make_query_set(2000, 2000, seed=77)
    t0, t1 = kw.get("t0", 0.0), kw.get("t1", 1.0)
    et, es, d = kw.get("eps_time", 1e-4), kw.get("eps_space", 1e-6), kw.get("depth", 32)
    for fam, x in ((1, qs.vf), (2, qs.ee)):
        res, _ = _device_batch(torch, fam, _native.prims_from_positions(x), t0, t1, _native.cert_config(et, es, d))
        _same(res, oracle.ref_cert(fam, oracle.prims_from_positions(x), t0, t1, et, es, d))


def test_golden_vectors_on_device(gpu):
    import torch

    g = np.load(oracle.HERE.parent / "tests" / "golden" / "ref_cert_golden.npz")
    keys = {}
    for i in range(len(g["family"])):
        keys.setdefault((int(g["family"][i]),) + tuple(g["params"][i]), []).append(i)
    for (fam, t0, t1, et, es, d), rows in keys.items():
        rows = np.array(rows)
        ids = np.stack([2 * rows + 10, 2 * rows + 11], axis=1)
        res, _ = _device_batch(torch, fam, _native.prims_from_positions(g["x"][rows], ids), t0, t1,
                               _native.cert_config(et, es, int(d)))
        for f in ("status", "toi_upper", "safe_margin_lb", "interval_t0", "interval_t1", "reason_code",
                  "witness_id_a", "witness_id_b", "covered_feature_mask", "next_refinement_mode"):
            assert np.array_equal(res[f].astype(np.float64), g[f][rows].astype(np.float64)), f


def test_invalid_primitives_are_flagged(gpu):
    import torch

    p = _native.prims_from_positions(np.random.default_rng(0).uniform(-1, 1, (64, 8, 3)))
    p["id_a"][5] = -3
    p["v"][9, 1]["position_t0"][0] = np.inf
    res, bad = _device_batch(torch, 1, p)
    assert bad == 2
    assert res["status"][5] == 255 and res["status"][9] == 255 and res["reason_code"][5] == 3
    with pytest.raises(_native.NativeError, match="ids are required"):
        _native.cert_ref_batch_host(1, p, 0.0, 1.0, _native.cert_config())


def test_module_batch_cross_check_and_queue(gpu):
    cpp = pytest.importorskip("p2cccd_cpp")
    m = cpp

    def v(i, a, b):
        t = m.LinearVertexTrajectory()
        t.feature_id, t.position_t0, t.position_t1 = i, a, b
        return t

    def pt(z0, z1, a=10, b=20):
        p = m.PointTriangleIntervalPrimitive()
        p.point_id, p.triangle_id = a, b
        p.point = v(a, [0.25, 0.25, z0], [0.25, 0.25, z1])
        p.triangle_v0, p.triangle_v1, p.triangle_v2 = v(101, [0, 0, 0], [0, 0, 0]), v(102, [1, 0, 0], [1, 0, 0]), \
            v(103, [0, 1, 0], [0, 1, 0])
        return p

    def ee(z0, z1, a=30, b=40):
        p = m.EdgeEdgeIntervalPrimitive()
        p.edge_a_id, p.edge_b_id = a, b
        p.edge_a0, p.edge_a1 = v(201, [-1, 0, 0], [-1, 0, 0]), v(202, [1, 0, 0], [1, 0, 0])
        p.edge_b0, p.edge_b1 = v(301, [0, -1, z0], [0, -1, z1]), v(302, [0, 1, z0], [0, 1, z1])
        return p

    cfg = m.CertificateEngineConfig()
    cfg.eps_time, cfg.eps_space, cfg.max_subdivision_depth = 1e-5, 1e-6, 32
    r = m.evaluate_point_triangle_batch_cuda([pt(1.0, -3.0), pt(2.0, 2.0, 11, 21)], 0.0, 1.0, cfg)
    assert r[0].status == m.CertificateStatus.COLLISION and abs(r[0].toi_upper - 0.25) <= 1e-12
    assert r[1].status == m.CertificateStatus.SEPARATION and r[1].safe_margin_lb > 1.9
    assert m.cross_check_cpu_cuda_exact([pt(1.0, -3.0), pt(2.0, 2.0, 11, 21)],
                                        [ee(1.0, -3.0), ee(2.0, 2.0, 31, 41)], 0.0, 1.0, cfg, 1e-9)

    def item(mask, wid):
        it = m.ExactWorkItem()
        it.work_item_id, it.parent_candidate_id, it.query_id = wid, 77, 1234
        it.interval_t0, it.interval_t1, it.feature_family_mask, it.priority_score = 0.0, 1.0, mask, 1.0
        return it

    qc, qs_, qu = m.ExactCertificateQuery(), m.ExactCertificateQuery(), m.ExactCertificateQuery()
    qc.work_item, qc.config, qc.point_triangle_primitives = item(1, 9001), cfg, [pt(1.0, -3.0)]
    qs_.work_item, qs_.config, qs_.edge_edge_primitives = item(2, 9002), cfg, [ee(2.0, 2.0)]
    qu.work_item, qu.config = item(1, 9003), cfg
    qcfg = m.ExactWorkQueueConfig()
    qcfg.first_event_id, qcfg.first_timestamp_us = 500, 800
    res = m.process_exact_work_queue_cpu([qc, qs_, qu], qcfg)
    assert res.processed_count == 3 and len(res.certificates) == 3 and len(res.audit_log) == 6
    assert [c.status for c in res.certificates] == [m.CertificateStatus.COLLISION, m.CertificateStatus.SEPARATION,
                                                     m.CertificateStatus.UNDECIDED]
    assert res.certificates[2].next_refinement_mode == m.CertificateRefinementMode.REQUEST_GEOMETRY
    assert res.audit_log[0].action == m.EXACT_AUDIT_DEQUEUED and res.audit_log[1].action == m.EXACT_AUDIT_COLLISION
    assert [r.event_id for r in res.audit_log] == list(range(500, 506))
    assert m.validate_exact_work_queue_coverage([qc, qs_, qu], res)


def test_engine_aggregation_matches_compiled_reference(gpu):
    cpp = pytest.importorskip("p2cccd_cpp")
    if not oracle.ref_available():
        pytest.skip("compiled reference not present")
    _engine_vs_reference(cpp)


def _engine_vs_reference(m):
    rng = np.random.default_rng(11)
    qs = workloads.make_query_set(120, 120, seed=8)
    cfg = m.CertificateEngineConfig()

    def to_prims(x, cls):
        out = []
        for k, q in enumerate(x):
            p = cls()
            trajs = []
            for j in range(4):
                t = m.LinearVertexTrajectory()
                t.feature_id, t.position_t0, t.position_t1 = j, list(q[j]), list(q[4 + j])
                trajs.append(t)
            if cls is m.PointTriangleIntervalPrimitive:
                p.point_id, p.triangle_id = 2 * k, 2 * k + 1
                p.point, p.triangle_v0, p.triangle_v1, p.triangle_v2 = trajs
            else:
                p.edge_a_id, p.edge_b_id = 2 * k, 2 * k + 1
                p.edge_a0, p.edge_a1, p.edge_b0, p.edge_b1 = trajs
            out.append(p)
        return out

    for trial in range(40):
        npt, nee = rng.integers(0, 4), rng.integers(0, 4)
        ia, ib = rng.integers(0, 120, npt), rng.integers(0, 120, nee)
        mask = int(rng.integers(1, 4))
        q = m.ExactCertificateQuery()
        q.work_item.work_item_id, q.work_item.parent_candidate_id, q.work_item.query_id = 1, 1, 1
        q.work_item.feature_family_mask = mask
        q.config = cfg
        q.point_triangle_primitives = to_prims(qs.vf[ia], m.PointTriangleIntervalPrimitive)
        q.edge_edge_primitives = to_prims(qs.ee[ib], m.EdgeEdgeIntervalPrimitive)
        got = m.evaluate_certificate_query_cpu(q)
        ref = oracle.ref_engine_compiled(mask, oracle.prims_from_positions(qs.vf[ia].reshape(-1, 8, 3),
                                                                           ids=np.stack([2 * np.arange(npt), 2 * np.arange(npt) + 1], 1)),
                                         oracle.prims_from_positions(qs.ee[ib].reshape(-1, 8, 3),
                                                                     ids=np.stack([2 * np.arange(nee), 2 * np.arange(nee) + 1], 1)))
        assert int(got.status) == int(ref[0]) and got.toi_upper == ref[1] and got.safe_margin_lb == ref[2]
        assert int(got.reason_code) == int(ref[7]) and int(got.next_refinement_mode) == int(ref[8])
        assert got.covered_feature_mask == int(ref[6])


@pytest.mark.parametrize("kw", [dict(), dict(max_depth=6), dict(eps_time=1e-8, max_depth=40)])
def test_cuda_twin_walk_matches_oracle(gpu, kw):
    # P2C_CERT_CUDA_TWIN: the walk of the reference's CUDA batch twin
    # (certificate_cuda.cu:286-346), UNDECIDED on the first unsplittable
    # interval; light and wave passes, node counts included.
    import torch

    et, d = kw.get("eps_time", 1e-4), kw.get("max_depth", 32)
    qs = workloads.make_query_set(3000, 3000, seed=99)
    for fam, x in ((1, qs.vf), (2, qs.ee)):
        cfg = _native.cert_config(et, 1e-6, d, cuda_twin=True)
        res, bad = _device_batch(torch, fam, _native.prims_from_positions(x), cfg=cfg)
        ref = oracle.ref_cert_twin(fam, oracle.prims_from_positions(x), eps_time=et, max_depth=d)
        assert bad == 0
        _same(res, ref)


def test_module_batch_is_the_cuda_twin(gpu):
    # evaluate_*_batch_cuda follows the CUDA twin, so on a pair where an
    # unsplittable interval precedes the first contact it answers UNDECIDED
    # and the CPU/CUDA cross-check reports the reference's status mismatch.
    m = pytest.importorskip("p2cccd_cpp")
    qs = workloads.make_query_set(1500, 1500, seed=99)
    x = qs.vf
    op = oracle.prims_from_positions(x)
    cpu, tw = oracle.ref_cert(1, op), oracle.ref_cert_twin(1, op)
    idx = np.nonzero((cpu["status"] == 0) & (tw["status"] == 2))[0][:4]
    assert idx.size > 0

    def v(fid, a, b):
        t = m.LinearVertexTrajectory()
        t.feature_id, t.position_t0, t.position_t1 = fid, list(a), list(b)
        return t

    prims = []
    for i in idx:
        p = m.PointTriangleIntervalPrimitive()
        p.point_id, p.triangle_id = int(i), int(i) + 100000
        p.point = v(0, x[i, 0], x[i, 4])
        p.triangle_v0, p.triangle_v1, p.triangle_v2 = (v(k, x[i, k], x[i, 4 + k]) for k in (1, 2, 3))
        prims.append(p)
    cfg = m.CertificateEngineConfig()
    cfg.eps_time, cfg.eps_space, cfg.max_subdivision_depth = 1e-4, 1e-6, 32
    r = m.evaluate_point_triangle_batch_cuda(prims, 0.0, 1.0, cfg)
    for k, i in enumerate(idx):
        assert r[k].status == m.CertificateStatus.UNDECIDED
        assert (r[k].interval_t0, r[k].interval_t1) == (tw["interval_t0"][i], tw["interval_t1"][i])
        assert r[k].toi_upper == tw["toi_upper"][i]
    with pytest.raises(Exception, match="status mismatch"):
        m.cross_check_cpu_cuda_exact(prims, [], 0.0, 1.0, cfg, 1e-9)


@pytest.mark.parametrize("kw", [dict(), dict(eps_time=1e-4, max_depth=12), dict(eps_time=1e-9, max_depth=40)])
def test_python_twin_walk_matches_oracle(gpu, kw):
    # P2C_CERT_PYTHON_TWIN: the reference's Python twin (pure_exact_cpu.py:
    # 98-308) against its restatement: status, TOI, margin, evaluations and
    # deepest depth, through the light and wave passes.
    import torch

    et, es, d = kw.get("eps_time", 1e-5), 1e-8, kw.get("max_depth", 32)
    qs = workloads.make_query_set(3000, 3000, seed=41)
    for fam, x in ((1, qs.vf), (2, qs.ee)):
        cfg = _native.cert_config(et, es, d)
        cfg["flags"] = 2
        res, bad = _device_batch(torch, fam, _native.prims_from_positions(x), cfg=cfg)
        ref = oracle.ref_cert_py(fam, oracle.prims_from_positions(x), eps_time=et, eps_space=es, max_depth=d)
        assert bad == 0
        _same(res, ref, ("status", "toi_upper", "safe_margin_lb", "nodes"))
        assert np.array_equal(res["pad"][:, 0].astype(np.int64), ref["reason_code"].astype(np.int64))


def test_external_ccd_query_module(gpu):
    # p2cccd.pure_exact mirrors evaluate_external_ccd_query: the reference's
    # checks (test_pure_exact_cpu_baseline.py:57-70, :101-108) and the closed
    # forms of the restatement's tests, through one batched launch.
    from types import SimpleNamespace as Q

    from p2cccd import pure_exact as pe

    tri = ((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0))

    def vf(z0, z1, qid):
        return Q(query_id=qid, family="vf", vertices_t0=((0.25, 0.25, z0),) + tri,
                 vertices_t1=((0.25, 0.25, z1),) + tri, ground_truth_collides=None)

    def ee(z0, z1, qid):
        return Q(query_id=qid, family="ee", vertices_t0=((-1.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, -1.0, z0), (0.0, 1.0, z0)),
                 vertices_t1=((-1.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, -1.0, z1), (0.0, 1.0, z1)),
                 ground_truth_collides=True)

    cfg = pe.PureExactCPUConfig(eps_time=1.0e-6, eps_space=1.0e-8, max_subdivision_depth=28)
    r = pe.evaluate_external_ccd_queries([vf(1.0, -1.0, 1), ee(1.0, -1.0, 2), vf(2.0, 2.0, 3), vf(1.0, -3.0, 4),
                                          vf(5e-9, -5e-9, 5)], cfg)
    assert [x.query_id for x in r] == [1, 2, 3, 4, 5]
    assert r[0].predicted_collision and r[0].toi_upper <= 0.5 and r[0].family == "point_triangle"
    assert r[1].predicted_collision and r[1].toi_upper <= 0.5 and r[1].family == "edge_edge"
    assert not r[2].predicted_collision and r[2].status == "separation"
    assert (r[0].toi_upper, r[0].exact_evals, r[0].max_depth) == (0.5, 3, 0)
    assert (r[3].status, r[3].toi_upper, r[3].exact_evals, r[3].max_depth) == ("collision", 0.25, 6, 1)
    assert (r[2].toi_upper, r[2].safe_margin_lb, r[2].exact_evals) == (1.0, 2.0 - 1.0e-8, 3)
    assert r[4].toi_upper == 0.5
    with pytest.raises(ValueError, match="eps_time"):
        pe.evaluate_external_ccd_query(vf(1.0, -1.0, 1), pe.PureExactCPUConfig(eps_time=0.0))


def test_python_twin_rules_closed_forms_on_device(gpu):
    # The two Python-twin rules on the device walk (P2C_CERT_PYTHON_TWIN),
    # against closed forms: the segment-segment guard abs(denom) > 1e-14
    # (nearly parallel crossing segments: the s = 0 branch's distance
    # 1.000297e-3, not the true 1e-3) and the motion radius 0.5 * (t1 - t0) *
    # (point + largest triangle-vertex displacement): margin 2.1 - 0.4 - eps.
    import torch
    from test_oracle_ref_cert import _ee_crossing, _pt_radius_case, ee_guard_closed_form

    cfg = _native.cert_config(1e-5, 1e-8, 32)
    cfg["flags"] = 2
    p, (L, h, z) = _ee_crossing()
    res, bad = _device_batch(torch, 2, p.view(_native.PRIM_DTYPE), cfg=cfg)
    assert bad == 0 and res["status"][0] == 1
    assert abs(res["safe_margin_lb"][0] + 1e-8 - ee_guard_closed_form(L, h, z)) <= 1e-15
    res, bad = _device_batch(torch, 1, _pt_radius_case().view(_native.PRIM_DTYPE), cfg=cfg)
    assert bad == 0 and res["status"][0] == 1 and res["nodes"][0] == 1
    assert abs(res["safe_margin_lb"][0] - (2.1 - 0.4 - 1e-8)) <= 4e-16
