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
