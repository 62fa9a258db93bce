"""Host-side checks of the batched broad phase's C-ABI: workspace sizes and
the refusals that come before any device work (no kernel is launched, the
scene addresses are never dereferenced on the host)."""
import pytest

from p2cccd import _native
from p2cccd.broadphase import _Mesh, _Scene

lib = _native.lib


def _scene(nv=10, nf=4, ne=9, nobj=1, ws=1 << 20, th=0.0, at=0):
    m = _Mesh(8, 8, 8, 8, 8, nv, nf, ne, nobj, 0, None, 0)
    return _Scene(m, 16, 16, th, 4096 + at * 8192, ws)


def _count(arr, n, slots=0, ws_bytes=1 << 40):
    offsets = (_native.ctypes.c_uint64 * 8)()
    return _native.call("p2c_bp_batch_count", arr, n, slots, 1 << 12, ws_bytes, _native.ctypes.addressof(offsets),
                        None)


def test_workspace_sizes():
    arr = (_Scene * 2)(_scene(), _scene(nv=5000, nf=10000, ne=15000, at=1))
    b0 = lib.p2c_bp_batch_workspace_bytes(arr, 2, 0)
    b16 = lib.p2c_bp_batch_workspace_bytes(arr, 2, 16)
    assert b16 - b0 >= 4 * 16 * (10 + 9 + 5000 + 15000) - 512  # the hit cache, 256-byte aligned regions
    assert lib.p2c_bp_batch_workspace_bytes(arr, 0, 0) == 0
    # 4 key arrays of whole 4096-key tiles: one tile for the first scene, ceil(30000 / 4096) for the second
    assert b0 >= 16 * 4096 * (1 + 8)
    # a resident scene keeps no sort keys, counts or offsets of its own
    assert 0 < lib.p2c_bp_scene_workspace_bytes(5000, 10000, 15000, 1) < \
        lib.p2c_bp_workspace_bytes(5000, 10000, 15000, 1)


@pytest.mark.parametrize("edit,text", [
    (dict(n=0), "1 to 1024 scenes"),
    (dict(slots=5000), "hit_cache_slots must be in"),
    (dict(scene=dict(nobj=0)), "object count must be in"),
    (dict(scene=dict(th=-1.0)), "thickness must be finite"),
    (dict(scene=dict(ws=64)), "scene workspace is too small"),
    (dict(ws_bytes=1024), "batch workspace is too small"),
    (dict(dup=True), "appears twice"),
])
def test_refusals_before_device_work(edit, text):
    a = _scene(**edit.get("scene", {}))
    b = _scene(nv=7, nf=2, ne=5, at=0 if edit.get("dup") else 1)
    arr = (_Scene * 2)(a, b)
    with pytest.raises(_native.NativeError, match=text):
        _count(arr, edit.get("n", 2), edit.get("slots", 0), edit.get("ws_bytes", 1 << 40))
