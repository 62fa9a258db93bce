"""gfx950 proxy-scene candidates against the reference's compiled CPU path
(golden vectors, tests/golden/proxy_golden.npz, including the KAT scene of
src/tests/cpp/test_rt_candidate.cpp:97-150) and against the oracle on a
larger seeded scene: raw hits identical field by field and in order,
compact candidates identical in key order, density identical."""
import numpy as np
import pytest

import oracle
from p2cccd import proxy

from test_oracle_proxy import golden_scenes

pytestmark = pytest.mark.gpu

RAW_FIELDS = ("query_id", "pair_key", "proxy_a_index", "proxy_b_index", "slab_id", "rt_hit_count", "flags",
              "motion_bound")
CAND_FIELDS = ("candidate_id", "query_id", "slab_id", "object_a_id", "patch_a_id", "object_b_id", "patch_b_id",
               "proxy_type_a", "proxy_type_b", "rt_hit_count", "motion_bound", "flags")


def _check(res, raw, cand, dens):
    r, c = res["raw_buffer"], res["candidates"]
    assert len(r) == len(raw) and len(c) == len(cand)
    for f in RAW_FIELDS:
        assert np.array_equal(r[f], raw[f]), f
    for f in CAND_FIELDS:
        assert np.array_equal(c[f].astype(np.uint64) if f != "motion_bound" else c[f],
                              cand[f].astype(np.uint64) if f != "motion_bound" else cand[f]), f
    for f in oracle.DENSITY.names:
        assert res["density"][f] == dens[f], f


def test_golden_scenes(gpu):
    n = 0
    for q, scene, raw, cand, dens in golden_scenes():
        _check(proxy.generate_candidates_for_proxy_scene(scene, q), raw, cand, dens)
        n += 1
    assert n == 7


def test_large_scene_against_oracle(gpu):
    rng = np.random.default_rng(31)
    s = oracle.random_proxy_scene(rng, 6000, 6, 3, extent=1.0, size=0.12)
    s["patch_id"] = (s["patch_id"] % 3).astype(np.uint32)
    res = proxy.generate_candidates_for_proxy_scene(s, 77)
    raw, cand, dens = oracle.proxy_generate(77, s)
    assert len(raw) > 1000 and len(cand) < len(raw)
    _check(res, raw, cand, dens)


def _records(objs, fields):
    return {f: np.array([np.array(getattr(o, f)) if f == "motion_bound" else
                         (int(getattr(o, f)) if not isinstance(getattr(o, f), float) else getattr(o, f))
                         for o in objs]) for f in fields}


def test_pybind_generate_matches_golden(gpu):
    """p2cccd_cpp.generate_candidates_for_proxy_scene (the reference's pybind
    entry) on the device path: identical to the compiled reference's output."""
    import p2cccd_cpp as cpp
    from test_proxy import cpp_scene
    for q, scene, raw, cand, dens in golden_scenes():
        for backend in ("hip_exact", "cpu_reference", "optix"):
            res = cpp.generate_candidates_for_proxy_scene(cpp_scene(q, scene), backend_name=backend)
            assert res.backend_name == "hip_exact"
            r = res.raw_buffer.hits
            c = res.candidates
            assert len(r) == len(raw) and len(c) == len(cand)
            if len(r):
                got = _records(r, RAW_FIELDS)
                for f in RAW_FIELDS:
                    assert np.array_equal(got[f].astype(raw[f].dtype), raw[f]), f
            if len(c):
                got = _records(c, CAND_FIELDS)
                for f in CAND_FIELDS:
                    assert np.array_equal(got[f].astype(cand[f].dtype), cand[f]), f
            for f in oracle.DENSITY.names:
                if f in ("timing", "backend_name", "query_id", "schema_version"):
                    continue
                assert getattr(res.density, f) == dens[f], f


def test_pybind_compact_reports_first_error_in_buffer_order(gpu):
    import p2cccd_cpp as cpp
    from test_proxy import cpp_scene
    q, scene, raw, cand, dens = list(golden_scenes())[5]
    s = cpp_scene(q, scene)
    buf = cpp.generate_raw_candidates_cpu(s)
    assert len(buf.hits) == len(raw) > 10
    out = cpp.compact_raw_candidates(s, buf)
    assert [int(c.candidate_id) for c in out] == list(range(1, len(cand) + 1))

    def broken(edits):
        b = cpp.RawCandidateBuffer()
        hits = list(cpp.generate_raw_candidates_cpu(s).hits)
        for i, fn in edits:
            fn(hits[i])
        b.hits = hits
        return b

    def setattr_(f, v):
        return lambda h: setattr(h, f, v)

    cases = [
        ([(3, setattr_("flags", 0))], "not marked valid"),
        ([(3, setattr_("query_id", 0))], "query_id is required"),
        ([(3, setattr_("query_id", q + 1))], "must match ProxyScene.query_id"),
        ([(3, setattr_("proxy_b_index", 10 ** 6))], "proxy index is out of range"),
        ([(3, setattr_("pair_key", 5))], "pair_key does not match"),
        ([(3, setattr_("motion_bound", [0.0, -1.0, 0.0, 0.0]))], "finite and non-negative"),
        # two broken hits: the earlier one in buffer order is reported
        ([(7, setattr_("flags", 0)), (2, setattr_("query_id", 0))], "query_id is required"),
        ([(2, setattr_("flags", 0)), (7, setattr_("query_id", 0))], "not marked valid"),
    ]
    for edits, msg in cases:
        with pytest.raises(ValueError, match=msg):
            cpp.compact_raw_candidates(s, broken(edits))
    # rt_hit_count overflow: two hits of one key summing past 2^32 - 1
    hits = list(cpp.generate_raw_candidates_cpu(s).hits)
    keys = {}
    pair = None
    for i, h in enumerate(hits):
        a, b = int(h.proxy_a_index), int(h.proxy_b_index)
        k = (int(scene["slab_id"][a]),) + tuple(sorted([(int(scene["object_id"][a]), int(scene["patch_id"][a]),
                                                         int(scene["proxy_type"][a])),
                                                        (int(scene["object_id"][b]), int(scene["patch_id"][b]),
                                                         int(scene["proxy_type"][b]))]))
        if k in keys:
            pair = (keys[k], i)
            break
        keys[k] = i
    assert pair is not None
    hits[pair[0]].rt_hit_count = 0xFFFFFFFF
    b = cpp.RawCandidateBuffer()
    b.hits = hits
    with pytest.raises(ValueError, match="rt_hit_count overflow during compaction"):
        cpp.compact_raw_candidates(s, b)


@pytest.mark.parametrize("n", [1, 5, 1023, 16384, 16385, 40000, 65536, 65537, 300001])
def test_exclusive_scan_both_forms(gpu, n):
    # Up to 65,536 items the scan runs in one workgroup, above in three
    # phases; both write the u64 exclusive prefix and the total at out[n].
    import torch
    from p2cccd import _native as nat  # p2c_exclusive_scan_u32 is declared by p2cccd.proxy

    rng = np.random.default_rng(n)
    vals = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
    vals[rng.random(n) < 0.3] = 0
    src = torch.from_numpy(vals.view(np.int32)).to("cuda")
    out = torch.full((n + 1,), -1, dtype=torch.int64, device="cuda")
    part = torch.empty(n // 1024 + 2, dtype=torch.int64, device="cuda")
    nat.call("p2c_exclusive_scan_u32", src.data_ptr(), n, out.data_ptr(), part.data_ptr(),
             nat.stream_handle(torch.cuda.current_stream()))
    torch.cuda.synchronize()
    want = np.zeros(n + 1, dtype=np.uint64)
    np.cumsum(vals.astype(np.uint64), out=want[1:])
    assert np.array_equal(out.cpu().numpy().view(np.uint64), want)
