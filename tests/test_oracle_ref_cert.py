"""The reference-parity oracle (oracle/ref_cert.c) pinned against the
reference's own compiled certificate: golden vectors, the reference's KATs
(src/tests/cpp/test_certificate_engine.cpp:97-158), closed forms, and a
planted mistake that the golden vectors must catch."""
import numpy as np
import pytest

import oracle
from p2cccd import workloads

GOLDEN = oracle.HERE.parent / "tests" / "golden" / "ref_cert_golden.npz"
FIELDS = [f for f in oracle.RESULT.names if f != "nodes"]


def _golden():
    g = np.load(GOLDEN)
    return g


def _run_rows(fn, g):
    out = np.zeros(len(g["family"]), oracle.RESULT)
    for i in range(len(g["family"])):
        t0, t1, et, es, d = g["params"][i]
        p = oracle.prims_from_positions(g["x"][i][None], ids=np.array([[2 * i + 10, 2 * i + 11]]))
        out[i] = fn(int(g["family"][i]), p, t0, t1, et, es, int(d))[0]
    return out


def test_oracle_matches_golden_vectors_bit_exact():
    g = _golden()
    out = _run_rows(oracle.ref_cert, g)
    for f in FIELDS:
        assert np.array_equal(out[f], g[f]), f
    # the fixture exercises every outcome
    assert set(np.unique(g["status"])) == {0, 1, 2}


def test_reference_kats():
    cfg = dict(eps_time=1e-5, eps_space=1e-6, max_depth=32)
    pt = oracle.prims_from_positions(np.array([_kat_pt(1.0, -3.0), _kat_pt(2.0, 2.0)]),
                                     ids=np.array([[10, 20], [10, 20]]))
    r = oracle.ref_cert(1, pt, **cfg)
    assert r["status"][0] == 0 and abs(r["toi_upper"][0] - 0.25) <= 1e-12
    assert r["witness_family"][0] == 1 and (r["witness_id_a"][0], r["witness_id_b"][0]) == (10, 20)
    assert r["status"][1] == 1 and r["covered_feature_mask"][1] == 1 and r["safe_margin_lb"][1] > 1.9
    ee = oracle.prims_from_positions(np.array([_kat_ee(1.0, -3.0), _kat_ee(2.0, 2.0)]),
                                     ids=np.array([[30, 40], [30, 40]]))
    r = oracle.ref_cert(2, ee, **cfg)
    assert r["status"][0] == 0 and abs(r["toi_upper"][0] - 0.25) <= 1e-12
    assert (r["witness_id_a"][0], r["witness_id_b"][0]) == (30, 40)
    assert r["status"][1] == 1 and r["covered_feature_mask"][1] == 2 and r["safe_margin_lb"][1] > 1.9


def test_closed_form_dyadic_crossing():
    # A point crossing a static plane at t* = z0 / (z0 - z1): when t* is a
    # dyadic midpoint the first sample <= eps_space is exactly t*.
    for z0, z1, tstar in [(1.0, -1.0, 0.5), (1.0, -3.0, 0.25), (3.0, -1.0, 0.75), (1.0, -7.0, 0.125)]:
        p = oracle.prims_from_positions(np.array([_kat_pt(z0, z1)]))
        r = oracle.ref_cert(1, p, eps_time=1e-9, eps_space=1e-12, max_depth=40)
        assert r["status"][0] == 0 and r["toi_upper"][0] == tstar
    # a static separation of h has margin exactly h - eps on the first node
    for h in (0.5, 2.0):
        r = oracle.ref_cert(1, oracle.prims_from_positions(np.array([_kat_pt(h, h)])), eps_space=1e-6)
        assert r["status"][0] == 1 and r["safe_margin_lb"][0] == h - 1e-6 and r["nodes"][0] == 1


@pytest.mark.parametrize("mutation", [1, 2, 3, 4])
def test_planted_mistake_is_caught_by_golden(mutation):
    # Each mutation is a plausible restatement error planted in the oracle
    # (subdivision order, dropped motion term, re-associated lerp, margin
    # offset); the reference's golden vectors must reject every one.
    import ctypes

    flag = ctypes.c_int.in_dll(oracle.lib(), "orc_ref_mutation")
    g = _golden()
    try:
        flag.value = mutation
        out = _run_rows(oracle.ref_cert, g)
    finally:
        flag.value = 0
    assert any(not np.array_equal(out[f], g[f]) for f in FIELDS)


@pytest.mark.skipif(not oracle.ref_available(), reason="compiled reference not present")
def test_oracle_matches_compiled_reference_random():
    qs = workloads.make_query_set(1500, 1500, seed=99)
    for fam, x in ((1, qs.vf), (2, qs.ee)):
        p = oracle.prims_from_positions(x)
        for kw in (dict(), dict(eps_time=1e-5), dict(t0=0.1, t1=0.9, max_depth=6)):
            a = oracle.ref_cert(fam, p, **kw)
            b = oracle.ref_cert_compiled(fam, p, **kw)
            for f in FIELDS:
                assert np.array_equal(a[f], b[f]), (fam, kw, f)


def _kat_pt(z0, z1):
    x = np.zeros((8, 3))
    x[0], x[4] = [0.25, 0.25, z0], [0.25, 0.25, z1]
    x[1] = x[5] = [0, 0, 0]
    x[2] = x[6] = [1, 0, 0]
    x[3] = x[7] = [0, 1, 0]
    return x


def _kat_ee(z0, z1):
    x = np.zeros((8, 3))
    x[0] = x[4] = [-1, 0, 0]
    x[1] = x[5] = [1, 0, 0]
    x[2], x[6] = [0, -1, z0], [0, -1, z1]
    x[3], x[7] = [0, 1, z0], [0, 1, z1]
    return x


@pytest.mark.skipif(not oracle.ref_available(), reason="compiled reference not present")
@pytest.mark.parametrize("kw", [dict(), dict(max_depth=6), dict(eps_time=1e-2), dict(t0=0.1, t1=0.9, max_depth=6)])
def test_cuda_twin_walk_against_compiled_engine(kw):
    # The CUDA twin (certificate_cuda.cu:286-346) cannot be compiled here; its
    # restatement is pinned through what ties it to the compiled CPU engine:
    # the two walks visit the same intervals in the same order and differ
    # only in what an interval that may not be split does (the twin stops on
    # it, the engine notes it and searches on).
    qs = workloads.make_query_set(1500, 1500, seed=99)
    T0, T1 = kw.get("t0", 0.0), kw.get("t1", 1.0)
    depth_cap, eps_time = kw.get("max_depth", 32), kw.get("eps_time", 1e-4)
    turned_total = 0
    for fam, x in ((1, qs.vf), (2, qs.ee)):
        p = oracle.prims_from_positions(x)
        cpu = oracle.ref_cert_compiled(fam, p, **kw)
        own = oracle.ref_cert(fam, p, **kw)  # the restated engine, for node counts
        tw = oracle.ref_cert_twin(fam, p, **kw)
        # separation: no interval touched or stopped either walk; every field equal
        sep = cpu["status"] == 1
        assert (tw["status"][sep] == 1).all()
        for f in FIELDS:
            assert np.array_equal(tw[f][sep], cpu[f][sep]), (fam, f)
        assert np.array_equal(tw["nodes"][sep], own["nodes"][sep])
        # a collision is the twin's too unless an unsplittable interval came first
        col = cpu["status"] == 0
        kept = col & (tw["status"] == 0)
        for f in FIELDS:
            assert np.array_equal(tw[f][kept], cpu[f][kept]), (fam, f)
        assert np.array_equal(tw["nodes"][kept], own["nodes"][kept])
        turned = col & (tw["status"] != 0)
        assert (tw["status"][turned] == 2).all()
        # that interval lies left of the collision (disjoint nodes: pre-order is left to right)
        assert (tw["interval_t1"][turned] <= cpu["interval_t0"][turned]).all()
        # the twin is never undecided where the engine separates; its undecided
        # results carry the engine's reason and refinement mode
        und = tw["status"] == 2
        assert not (und & sep).any()
        assert (tw["reason_code"][und] == 2).all() and (tw["next_refinement_mode"][und] == 1).all()
        for i in np.nonzero(und)[0]:
            a, b = tw["interval_t0"][i], tw["interval_t1"][i]
            # the compiled engine on that interval alone, with no split allowed,
            # finds neither a contact nor a separation
            leaf = oracle.ref_cert_compiled(fam, p[i:i + 1], t0=a, t1=b, eps_time=eps_time, max_depth=0)[0]
            assert leaf["status"] == 2, (fam, i)
            # and it may not be split: at the depth cap, or no wider than eps_time
            depth = int(round(np.log2((T1 - T0) / (b - a))))
            assert depth >= depth_cap or b - a <= eps_time, (fam, i, depth)
        turned_total += int(turned.sum())
    if not kw or "eps_time" in kw:
        assert turned_total > 0  # the workload holds cases where the two walks disagree


def _pt_z(z0, z1):
    # the reference's PT fixture (src/tests/python/test_pure_exact_cpu_baseline.py:26-37)
    x = np.zeros((1, 8, 3))
    x[0, 0], x[0, 4] = (0.25, 0.25, z0), (0.25, 0.25, z1)
    for k, v in ((1, (0, 0, 0)), (2, (1, 0, 0)), (3, (0, 1, 0))):
        x[0, k] = x[0, 4 + k] = v
    return oracle.prims_from_positions(x)


def _ee_z(z0, z1):
    # the reference's EE fixture (:40-51)
    x = np.zeros((1, 8, 3))
    x[0, 0] = x[0, 4] = (-1, 0, 0)
    x[0, 1] = x[0, 5] = (1, 0, 0)
    x[0, 2], x[0, 6] = (0, -1, z0), (0, -1, z1)
    x[0, 3], x[0, 7] = (0, 1, z0), (0, 1, z1)
    return oracle.prims_from_positions(x)


def test_python_twin_restatement_kats():
    # The reference's own checks (test_pure_exact_cpu_baseline.py:57-70),
    # then closed forms on the same fixtures, where the distance is |z(t)|
    # and z(t) = (1 - t) z0 + t z1 is exact at the sampled times.
    kw = dict(eps_time=1.0e-6, eps_space=1.0e-8, max_depth=28)
    pt = oracle.ref_cert_py(1, _pt_z(1.0, -1.0), **kw)[0]
    ee = oracle.ref_cert_py(2, _ee_z(1.0, -1.0), **kw)[0]
    sep = oracle.ref_cert_py(1, _pt_z(2.0, 2.0), **kw)[0]
    assert pt["status"] == 0 and pt["toi_upper"] <= 0.5
    assert ee["status"] == 0 and ee["toi_upper"] <= 0.5
    assert sep["status"] == 1
    # z 1 -> -1: the root's midpoint touches; 3 evaluations, depth 0
    assert (pt["toi_upper"], pt["nodes"], pt["reason_code"]) == (0.5, 1, 0)
    # z 1 -> -3: root [0,1] gives |z| = 1, 1, 3 and bound 1 - 0.5 * 4 < 0; its
    # left half touches at t = 0.25: 6 evaluations, depth 1
    r = oracle.ref_cert_py(1, _pt_z(1.0, -3.0), **kw)[0]
    assert (r["status"], r["toi_upper"], r["nodes"], r["reason_code"]) == (0, 0.25, 2, 1)
    # static at z = 2: one node, margin 2 - eps_space, TOI t1
    assert (sep["toi_upper"], sep["safe_margin_lb"], sep["nodes"]) == (1.0, 2.0 - 1.0e-8, 1)
    # contact time is the smallest sample, not the first within eps_space:
    # z 5e-9 -> -5e-9 touches at every sample, smallest (0) at the midpoint;
    # the C++ engine reports the first, t = 0
    r = oracle.ref_cert_py(1, _pt_z(5e-9, -5e-9), **kw)[0]
    assert r["status"] == 0 and r["toi_upper"] == 0.5
    e = oracle.ref_cert(1, _pt_z(5e-9, -5e-9), eps_time=1e-6, eps_space=1e-8, max_depth=28)[0]
    assert e["status"] == 0 and e["toi_upper"] == 0.0
    # an undecided walk reports t1 and margin 0, and its depth is the cap
    r = oracle.ref_cert_py(1, _pt_z(1.0, -1.0 + 1e-3), eps_time=1e-6, eps_space=1e-8, max_depth=2)[0]
    assert r["status"] in (0, 2)
    if r["status"] == 2:
        assert (r["toi_upper"], r["safe_margin_lb"], r["reason_code"]) == (1.0, 0.0, 2)


def test_python_twin_agrees_with_engine_decisions():
    # Same subdivision and tests as the C++ engine, up to the contact-time
    # rule and the rounding of the motion radius: the decisions of the two
    # agree on nearly every pair of the plain workload, and where both touch,
    # the twin's contact time lies in the engine's contact interval or later.
    qs = workloads.make_query_set(1500, 1500, seed=99)
    for fam, x in ((1, qs.vf), (2, qs.ee)):
        p = oracle.prims_from_positions(x)
        e = oracle.ref_cert(fam, p, eps_time=1e-5, eps_space=1e-8)
        y = oracle.ref_cert_py(fam, p, eps_time=1e-5, eps_space=1e-8)
        assert np.mean(e["status"] == y["status"]) > 0.98
        both = (e["status"] == 0) & (y["status"] == 0)
        assert (y["toi_upper"][both] >= e["interval_t0"][both]).all()


def _ee_crossing():
    """Static nearly parallel segments of length L = 2^-10 at height z = 1e-3
    whose projections cross at their midpoints, at angle h = 0.05: the
    denominator a e - b^2 = L^4 h^2 (2.27e-15) is below the twin's guard."""
    L, h, z = 2.0 ** -10, 0.05, 1e-3
    x = np.zeros((1, 8, 3))
    x[0, 0], x[0, 1] = (0, 0, 0), (L, 0, 0)
    x[0, 2], x[0, 3] = (0, -h * L / 2, z), (L, h * L / 2, z)
    x[0, 4:] = x[0, :4]
    return oracle.prims_from_positions(x), (L, h, z)


def ee_guard_closed_form(L, h, z) -> float:
    """Distance the twin's segment-segment rule gives when abs(denom) <= 1e-14
    (pure_exact_cpu.py:126): s = 0, t = f / e, the closest points p1 and
    p2 + t d2; in exact rational arithmetic, then a 50-digit square root."""
    import decimal
    from fractions import Fraction as Fr

    Lq, hq, zq = Fr(L), Fr(h), Fr(z)
    d2, r = (Lq, hq * Lq, Fr(0)), (Fr(0), hq * Lq / 2, -zq)
    t = sum(u * v for u, v in zip(d2, r)) / sum(c * c for c in d2)
    q = (t * Lq, -hq * Lq / 2 + t * hq * Lq, zq)
    d2sum = sum(c * c for c in q)
    decimal.getcontext().prec = 50
    return float(decimal.Decimal(d2sum.numerator).sqrt() / decimal.Decimal(d2sum.denominator).sqrt())


def test_python_twin_ee_guard_closed_form():
    # With the twin's guard abs(denom) > 1e-14 the crossing pair's distance is
    # the s = 0 branch's, 1.000297e-3; the engine's guard (denom != 0) solves
    # for the crossing and gets the true 1e-3.  Static pair: separation at the
    # root with margin = distance - eps_space.
    p, (L, h, z) = _ee_crossing()
    want = ee_guard_closed_form(L, h, z)
    tw = oracle.ref_cert_py(2, p, eps_time=1e-5, eps_space=1e-8, max_depth=32)[0]
    en = oracle.ref_cert(2, p, eps_time=1e-5, eps_space=1e-8)[0]
    assert tw["status"] == 1 and en["status"] == 1
    assert abs((tw["safe_margin_lb"] + 1e-8) - want) <= 1e-15
    assert abs(en["safe_margin_lb"] + 1e-8 - z) <= 1e-15 and abs(want - z) > 2e-7


def _pt_radius_case():
    """A point moving 0.6 along x at height 2 above the unit triangle, which
    drops 0.2 along z: distance at t = 0.5 is 2.1; point displacement 0.6,
    every triangle vertex 0.2."""
    x = np.zeros((1, 8, 3))
    x[0, 0], x[0, 4] = (0.1, 0.1, 2.0), (0.7, 0.1, 2.0)
    for k, v in ((1, (0, 0, 0)), (2, (1, 0, 0)), (3, (0, 1, 0))):
        x[0, k] = v
        x[0, 4 + k] = (v[0], v[1], -0.2)
    return oracle.prims_from_positions(x)


def test_python_twin_motion_radius_closed_form():
    # The twin's radius is 0.5 * (t1 - t0) * (point displacement + the largest
    # triangle-vertex displacement), the norms taken once over the whole step
    # (pure_exact_cpu.py:259-262): at the root 0.5 * (0.6 + 0.2) = 0.4, so
    # the margin is 2.1 - 0.4 - eps_space.  A sum over the triangle's
    # vertices (1.0 in place of 0.2) or the full length (no 0.5) would give
    # 1.5 or 1.3.  On dyadic intervals (every node of a [0, 1] walk) the
    # engine's per-vertex product 0.5 * len * |d| equals it bit for bit:
    # scaling by a power of two is exact, so both margins agree.
    p = _pt_radius_case()
    tw = oracle.ref_cert_py(1, p, eps_time=1e-5, eps_space=1e-8, max_depth=32)[0]
    en = oracle.ref_cert(1, p, eps_time=1e-5, eps_space=1e-8)[0]
    assert tw["status"] == 1 and tw["nodes"] == 1
    assert abs(tw["safe_margin_lb"] - (2.1 - 0.4 - 1e-8)) <= 4e-16
    assert tw["safe_margin_lb"] == en["safe_margin_lb"]
