"""External query batches and proxy-scene construction, host side (no kernel
is launched): the oracle's restatement of GenerateCandidatesForExternalBatch
and p2cccd_cpp.build_proxy_scene against the reference's own compiled code
(golden vectors, tests/golden/external_golden.npz, written by
tests/golden/make_external_golden.py), the reference's KATs
(src/tests/cpp/test_rt_candidate.cpp:53-98, 163-212, 276-283) and the
validation errors that the host raises before any device work."""
from pathlib import Path
from types import SimpleNamespace as NS

import numpy as np
import pytest

import oracle

cpp = pytest.importorskip("p2cccd_cpp")
G = np.load(Path(__file__).resolve().parent / "golden" / "external_golden.npz")


# ---- the restatement, pinned on the reference's outputs -----------------------
@pytest.mark.parametrize("k", range(int(G["e_ok_count"])))
def test_external_restatement_matches_reference(k):
    got, rid = oracle.external_batch(G[f"e{k}_q"])
    ref = G[f"e{k}_cand"]
    assert np.array_equal(rid, G[f"e{k}_rid"])
    assert len(got) == len(ref)
    for g, r in zip(got, ref):
        for f in ("candidate_id", "query_id", "slab_id", "object_a_id", "patch_a_id", "object_b_id", "patch_b_id"):
            assert g[f] == r[f], f
        assert np.float32(g["motion_bound"]) == r["motion_bound"][3] and (r["motion_bound"][:3] == 0).all()


@pytest.mark.parametrize("k", range(int(G["e_err_count"])))
def test_external_restatement_errors_match_reference(k):
    with pytest.raises(ValueError) as e:
        oracle.external_batch(G[f"x{k}_q"])
    assert str(e.value) == str(G[f"x{k}_msg"][0])


# ---- build_proxy_scene (host, bit-exact) -------------------------------------
def _build_input(query_id, objs, patches, segs):
    inp = cpp.ProxySceneBuildInput()
    inp.query_id = int(query_id)
    objects = []
    for o in objs:
        ob = cpp.ProxyObjectBuildInput()
        ob.object_id = int(o["object_id"])
        ob.proxy_type = cpp.ProxyType(int(o["proxy_type"])) if int(o["proxy_type"]) in (1, 2) else \
            cpp.ProxyType.UNKNOWN
        ob.slabs_per_motion_segment = int(o["slabs_per_motion_segment"])
        ob.eps_proxy = float(o["eps_proxy"])
        pats = []
        for p in patches[int(o["patch0"]):int(o["patch0"]) + int(o["n_patches"])]:
            pt = cpp.Patch()
            pt.patch_id, pt.local_center, pt.radius = int(p["patch_id"]), p["local_center"].tolist(), float(p["radius"])
            pats.append(pt)
        ob.patches = pats
        ms = []
        for s in segs[int(o["segment0"]):int(o["segment0"]) + int(o["n_segments"])]:
            m = cpp.MotionSegment()
            m.t0, m.t1 = float(s["t0"]), float(s["t1"])
            a, b = cpp.PoseSample(), cpp.PoseSample()
            a.translation, a.rotation_xyzw = s["translation0"].tolist(), s["rotation0"].tolist()
            b.translation, b.rotation_xyzw = s["translation1"].tolist(), s["rotation1"].tolist()
            m.pose_t0, m.pose_t1 = a, b
            ms.append(m)
        ob.motion_segments = ms
        objects.append(ob)
    inp.objects = objects
    return inp


def _flatten(p):
    mb = p.motion_bound
    return [p.proxy_id, p.object_id, p.patch_id, p.slab_id, p.motion_segment_id, int(p.proxy_type), p.t0, p.t1,
            *p.bounds.min, *p.bounds.max, *p.capsule.endpoint0, *p.capsule.endpoint1, p.capsule.radius, mb.patch_id,
            mb.t0, mb.t1, *mb.center_t0, *mb.center_t1, mb.translation_bound, mb.rotation_angle,
            mb.center_rotation_bound, mb.surface_rotation_bound, mb.radial_motion_bound, mb.conservative_radius]


@pytest.mark.parametrize("k", range(int(G["b_ok_count"])))
def test_build_proxy_scene_bit_exact(k):
    scene = cpp.build_proxy_scene(_build_input(42 + k, G[f"b{k}_o"], G[f"b{k}_p"], G[f"b{k}_s"]))
    got = np.array([_flatten(p) for p in scene.primitives], dtype=np.float64)
    ref = G[f"b{k}_prims"]
    assert scene.query_id == 42 + k and got.shape == ref.shape
    assert np.array_equal(got, ref)  # every bound, centre, angle and radius bit for bit
    assert cpp.validate_proxy_scene(scene)


@pytest.mark.parametrize("k", range(int(G["b_err_count"])))
def test_build_proxy_scene_errors_match_reference(k):
    with pytest.raises(ValueError) as e:
        cpp.build_proxy_scene(_build_input(7, G[f"y{k}_o"], G[f"y{k}_p"], G[f"y{k}_s"]))
    assert str(e.value) == str(G[f"y{k}_msg"][0])


def _kat_input():
    # MakeOverlappingSceneInput (test_rt_candidate.cpp:53-98)
    def patch(pid, x):
        p = cpp.Patch()
        p.patch_id, p.triangle_ids, p.triangle_count, p.area, p.local_center, p.radius = pid, [pid], 1, 1.0, [x, 0, 0], 0.2
        return p

    def static(x):
        m = cpp.MotionSegment()
        a = cpp.PoseSample()
        a.translation, a.rotation_xyzw = [x, 0.0, 0.0], [0.0, 0.0, 0.0, 1.0]
        m.t0, m.t1, m.pose_t0, m.pose_t1 = 0.0, 1.0, a, a
        return m

    objs = []
    for oid, typ, pid, x in ((10, cpp.ProxyType.SWEPT_AABB, 1, 0.0), (20, cpp.ProxyType.CAPSULE, 2, 0.25)):
        o = cpp.ProxyObjectBuildInput()
        o.object_id, o.proxy_type, o.patches, o.motion_segments = oid, typ, [patch(pid, x)], [static(0.0)]
        o.slabs_per_motion_segment, o.eps_proxy = 2, 0.05
        objs.append(o)
    inp = cpp.ProxySceneBuildInput()
    inp.query_id, inp.objects = 42, objs
    return inp


def test_build_proxy_scene_reference_kat():
    scene = cpp.build_proxy_scene(_kat_input())
    assert len(scene.primitives) == 4 and [p.slab_id for p in scene.primitives] == [0, 1, 0, 1]
    assert scene.primitives[0].bounds.min[0] == pytest.approx(-0.25) and scene.primitives[2].capsule.radius == 0.25
    if oracle.ref_available():  # the reference's own build of the same input
        q, arrays = oracle.ref_proxy_kat_scene()
        assert q == scene.query_id
        b = np.array([list(p.bounds.min) + list(p.bounds.max) for p in scene.primitives])
        assert np.array_equal(b, arrays["bounds"])
    bad = _kat_input()
    objs = bad.objects
    objs[0].proxy_type = cpp.ProxyType.UNKNOWN
    bad.objects = objs
    with pytest.raises(ValueError, match="proxy_type must be concrete"):
        cpp.build_proxy_scene(bad)
    bad = _kat_input()
    objs = bad.objects
    objs[1].slabs_per_motion_segment = 3
    bad.objects = objs
    with pytest.raises(ValueError, match="one global time interval"):
        cpp.build_proxy_scene(bad)


# ---- external batch: errors the host raises before any device work -----------
def _batch_from(q):
    items = []
    for r in q:
        items.append(NS(query_id=int(r["source_query_id"]), source_query_index=int(r["source_query_index"]),
                        family="vf" if r["family"] == 0 else "ee", vertices_t0=r["v0"].tolist(),
                        vertices_t1=r["v1"].tolist(),
                        ground_truth_collides=bool(r["ground_truth_collides"]) if r["has_ground_truth"] else None,
                        box_pair=tuple(int(v) for v in r["box_pair"]) if r["has_box_pair"] else None))
    return NS(queries=items)


@pytest.mark.parametrize("k", [0, 1])
def test_external_batch_host_errors_match_reference(k):
    with pytest.raises(ValueError) as e:
        cpp.generate_candidates_for_external_batch(_batch_from(G[f"x{k}_q"]))
    assert str(e.value) == str(G[f"x{k}_msg"][0])


def test_external_batch_argument_rules():
    q = _batch_from(G["e0_q"])
    with pytest.raises(ValueError, match="unsupported candidate backend"):
        cpp.generate_candidates_for_external_batch(q, backend_name="magic")
    q.queries[0].family = "xx"
    with pytest.raises(ValueError, match="unsupported external CCD query family: xx"):
        cpp.generate_candidates_for_external_batch(q)
    q = _batch_from(G["e0_q"])
    q.queries[1].box_pair = (1, 2, 3)
    with pytest.raises(ValueError, match="box_pair must contain exactly two object ids"):
        cpp.generate_candidates_for_external_batch(q)
    with pytest.raises(ValueError, match="at least one query"):
        cpp.generate_candidates_for_external_batch(NS(queries=[]))
