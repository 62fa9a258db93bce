"""Proxy-scene validation (host rules of src/cpp/rt_candidate/proxy_scene.cpp:
18-154) with the reference's messages, and the density arithmetic against the
oracle on the golden scenes.  No kernel is launched."""
import numpy as np
import pytest

import oracle
from p2cccd import proxy

from test_oracle_proxy import golden_scenes


def _scene():
    return oracle.random_proxy_scene(np.random.default_rng(1), 20, 2, 2)


@pytest.mark.parametrize("mutate,msg", [
    (lambda s: None, None),
    (lambda s: s["proxy_type"].__setitem__(0, 3), "unsupported proxy type"),
    (lambda s: s["t1"].__setitem__(0, s["t0"][0]), "finite t0 < t1"),
    (lambda s: s["bounds"].__setitem__((0, 0), np.nan), "non-finite value"),
    (lambda s: s["bounds"].__setitem__((0, 0), s["bounds"][0, 3] + 1), "min cannot exceed max"),
    (lambda s: s["t1"].__setitem__(np.flatnonzero(s["slab_id"] == s["slab_id"][0])[1], 0.75),
     "one global time interval"),
])
def test_validation_messages(mutate, msg):
    s = _scene()
    mutate(s)
    if msg is None:
        proxy.validate_proxy_scene(5, s)
    else:
        with pytest.raises(ValueError, match=msg):
            proxy.validate_proxy_scene(5, s)
    with pytest.raises(ValueError, match="query_id is required"):
        proxy.validate_proxy_scene(0, _scene())


def test_overlapping_slabs_are_rejected():
    s = _scene()
    s["slab_id"][:] = np.arange(20) % 2
    s["t0"][:] = np.where(s["slab_id"] == 0, 0.0, 0.25)
    s["t1"][:] = np.where(s["slab_id"] == 0, 0.5, 1.0)
    with pytest.raises(ValueError, match="must not overlap"):
        proxy.validate_proxy_scene(5, s)


def test_density_arithmetic_matches_oracle():
    for q, scene, raw, cand, dens in golden_scenes():
        c = np.zeros(len(cand), proxy.CANDIDATE_DTYPE)
        c["rt_hit_count"] = cand["rt_hit_count"]
        d = proxy.density(scene, len(raw), c)
        for f in oracle.DENSITY.names:
            assert d[f] == dens[f], f


def cpp_scene(q: int, scene: dict):
    """The scene as the reference's pybind objects (ProxyScene of
    ProxyPrimitive, proxy_id = index)."""
    import p2cccd_cpp as cpp
    s = cpp.ProxyScene()
    s.query_id = q
    prims = []
    mb_fields = ("translation_bound", "rotation_angle", "radial_motion_bound", "conservative_radius")
    for i in range(len(scene["t0"])):
        p = cpp.ProxyPrimitive()
        p.proxy_id = i
        p.object_id = int(scene["object_id"][i])
        p.patch_id = int(scene["patch_id"][i])
        p.slab_id = int(scene["slab_id"][i])
        p.proxy_type = cpp.ProxyType(int(scene["proxy_type"][i]))
        p.t0, p.t1 = float(scene["t0"][i]), float(scene["t1"][i])
        b = cpp.Aabb()
        b.min = [float(v) for v in scene["bounds"][i][:3]]
        b.max = [float(v) for v in scene["bounds"][i][3:]]
        p.bounds = b
        m = cpp.PatchMotionBound()
        for f, v in zip(mb_fields, scene["motion"][i]):
            setattr(m, f, float(v))
        p.motion_bound = m
        prims.append(p)
    s.primitives = prims
    return s


def test_pybind_validation_and_backend_names():
    """Host-side checks of the pybind surface: ValidateProxyScene messages
    (including proxy_id = index), backend-name parsing, and the messages the
    device compaction reports, before any kernel runs."""
    import p2cccd_cpp as cpp
    q, scene, *_ = next(golden_scenes())
    s = cpp_scene(q, scene)
    assert cpp.validate_proxy_scene(s)
    prims = list(s.primitives)
    prims[1].proxy_id = 7
    s.primitives = prims
    with pytest.raises(ValueError, match="proxy_id must match primitive array index"):
        cpp.validate_proxy_scene(s)
    s = cpp_scene(q, scene)
    s.query_id = 0
    with pytest.raises(ValueError, match="ProxyScene.query_id is required"):
        cpp.validate_proxy_scene(s)
    with pytest.raises(ValueError, match="unsupported candidate backend: bogus"):
        cpp.generate_candidates_for_proxy_scene(cpp_scene(q, scene), backend_name="bogus")
    msg = proxy.nat.lib.p2c_proxy_status_message
    assert msg(0) == b""
    assert msg(1 + ((5 << 4) | 9)) == b"raw candidate motion_bound must be finite and non-negative"
    assert msg(1 + ((0 << 4) | 10)) == b"candidate rt_hit_count overflow during compaction"


def test_raw_hit_workspace_size_and_validation():
    # Host-side checks of the C-ABI, before any device work: the workspace
    # holds 4 bytes per (column segment, row), at most 64 segments of at
    # least 1,024 columns; a short workspace is refused with a message.
    import ctypes

    from p2cccd import _native as nat

    ws = nat.lib.p2c_proxy_raw_workspace_bytes
    assert ws(1) == 4 and ws(1024) == 4 * 1024 and ws(1025) == 4 * 1025 * 2
    assert ws(64 * 1024) == 4 * 64 * 1024 * 64 and ws(10 ** 7) == 4 * 10 ** 7 * 64
    s = proxy.ProxySet(*([1] * len(proxy.FIELDS)), 5000, 7)
    rc = nat.lib.p2c_proxy_raw_count(ctypes.addressof(s), ctypes.c_void_p(8), ctypes.c_void_p(8), ws(5000) - 1, None)
    assert rc != 0 and b"workspace" in nat.lib.p2c_last_error()
    rc = nat.lib.p2c_proxy_raw_fill(ctypes.addressof(s), ctypes.c_void_p(8), None, ws(5000), ctypes.c_void_p(8), 1,
                                    None)
    assert rc != 0 and b"workspace" in nat.lib.p2c_last_error()
