"""CPU oracle for the proposal-to-certificate path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker / CPU baseline; the shipped library
never links or calls it.  See oracle.h for what each C file restates and how
each is pinned.

Two shared libraries are loaded lazily:
  liboracle.so        the C restatement (oracle/*.c), always built by build()
  _ref/libp2c_ref.so  the reference's own certificate_engine.cpp compiled in
                      place (oracle/Makefile `ref`); only exists where the
                      reference sources were present at build time.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libp2c_ref.so"

TRAJ = np.dtype([("feature_id", "<i8"), ("p0", "<f8", (3,)), ("p1", "<f8", (3,))])
PRIM = np.dtype([("id_a", "<i8"), ("id_b", "<i8"), ("v", TRAJ, (4,))])
RESULT = np.dtype(
    [
        ("status", "<i4"),
        ("covered_feature_mask", "<u4"),
        ("interval_t0", "<f8"),
        ("interval_t1", "<f8"),
        ("toi_upper", "<f8"),
        ("safe_margin_lb", "<f8"),
        ("witness_family", "<i4"),
        ("reason_code", "<i4"),
        ("witness_id_a", "<i8"),
        ("witness_id_b", "<i8"),
        ("next_refinement_mode", "<i4"),
        ("nodes", "<i4"),
    ]
)
TUV_RESULT = np.dtype([("hit", "<i4"), ("capped", "<i4"), ("toi", "<f8"), ("boxes", "<i8")])
assert PRIM.itemsize == 240 and RESULT.itemsize == 72 and TUV_RESULT.itemsize == 24

_lib = None
_ref = None
_vp, _i64, _i32, _f64 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise ImportError(f"{LIB} not built (make -C oracle)")
        _lib = ctypes.CDLL(str(LIB))
        _lib.orc_ref_eval_batch.argtypes = [_i32, _vp, _i64, _f64, _f64, _f64, _f64, _i32, _vp]
        _lib.orc_ref_eval_batch.restype = _i32
        _lib.orc_ref_eval_twin_batch.argtypes = [_i32, _vp, _i64, _f64, _f64, _f64, _f64, _i32, _vp]
        _lib.orc_ref_eval_twin_batch.restype = _i32
        _lib.orc_ref_eval_py_batch.argtypes = [_i32, _vp, _i64, _f64, _f64, _i32, _vp]
        _lib.orc_ref_eval_py_batch.restype = _i32
        _lib.orc_pt_distance.argtypes = [_vp, _f64]
        _lib.orc_pt_distance.restype = _f64
        _lib.orc_ee_distance.argtypes = [_vp, _f64]
        _lib.orc_ee_distance.restype = _f64
        _lib.orc_tiny_features.argtypes = [_i32, _vp, ctypes.c_uint64, _vp]
        _lib.orc_tiny_features.restype = None
        _lib.orc_tiny_predict.argtypes = [_vp, _vp]
        _lib.orc_tiny_predict.restype = _f64
        _lib.orc_tiny_score_batch.argtypes = [_i32, _i32, _vp, _vp, _vp, _i64, _vp, _vp]
        _lib.orc_tiny_score_batch.restype = None
        _lib.orc_proxy_generate.restype = _i32
        _lib.orc_unit_hash.argtypes = [ctypes.c_uint64]
        _lib.orc_unit_hash.restype = _f64
        if hasattr(_lib, "orc_tuv_eval_batch"):
            _lib.orc_tuv_eval_batch.argtypes = [_i32, _vp, _i64, _f64, _f64, _f64, _i64, _i64, _vp, _i32]
            _lib.orc_tuv_eval_batch.restype = _i32
            _lib.orc_tuv_params.argtypes = [_i32, _vp, _f64, _vp, _vp]
            _lib.orc_tuv_params.restype = None
    return _lib


def ref_available() -> bool:
    return REF_LIB.exists()


def ref_lib() -> ctypes.CDLL:
    global _ref
    if _ref is None:
        if not REF_LIB.exists():
            raise ImportError(f"{REF_LIB} not built (make -C oracle ref, needs the reference sources)")
        _ref = ctypes.CDLL(str(REF_LIB))
        _ref.ref_eval_batch.argtypes = [_i32, _vp, _i64, _f64, _f64, _f64, _f64, _i32, _vp]
        _ref.ref_eval_batch.restype = _i64
        _ref.ref_engine_eval.argtypes = [ctypes.c_uint32, _f64, _f64, _vp, _i64, _vp, _i64, _f64, _f64, _i32, _vp]
        _ref.ref_engine_eval.restype = _i32
    return _ref


def prims_from_positions(x: np.ndarray, ids: np.ndarray | None = None) -> np.ndarray:
    """(n, 8, 3) positions (4 vertices at t0 then at t1) -> PRIM records."""
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[0]
    out = np.zeros(n, PRIM)
    if ids is None:
        ids = np.stack([np.arange(n) * 2, np.arange(n) * 2 + 1], axis=1)
    out["id_a"], out["id_b"] = ids[:, 0], ids[:, 1]
    for k in range(4):
        out["v"][:, k]["feature_id"] = np.arange(n) * 4 + k
        out["v"][:, k]["p0"] = x[:, k]
        out["v"][:, k]["p1"] = x[:, 4 + k]
    return out


def ref_cert(family: int, prims: np.ndarray, t0=0.0, t1=1.0, eps_time=1.0e-4, eps_space=1.0e-6,
             max_depth=32) -> np.ndarray:
    """Restated reference certificate over a batch (oracle/ref_cert.c)."""
    prims = np.ascontiguousarray(prims, dtype=PRIM)
    out = np.zeros(prims.shape[0], RESULT)
    rc = lib().orc_ref_eval_batch(family, prims.ctypes.data, prims.shape[0], t0, t1, eps_time, eps_space,
                                  max_depth, out.ctypes.data)
    if rc != 0:
        raise ValueError(f"oracle rejected input (code {rc})")
    return out


def ref_cert_twin(family: int, prims: np.ndarray, t0=0.0, t1=1.0, eps_time=1.0e-4, eps_space=1.0e-6,
                  max_depth=32) -> np.ndarray:
    """Restated walk of the reference's CUDA batch twin (oracle/ref_cert.c,
    certificate_cuda.cu:286-346): UNDECIDED at the first interval that may not
    be split."""
    prims = np.ascontiguousarray(prims, dtype=PRIM)
    out = np.zeros(prims.shape[0], RESULT)
    rc = lib().orc_ref_eval_twin_batch(family, prims.ctypes.data, prims.shape[0], t0, t1, eps_time, eps_space,
                                       max_depth, out.ctypes.data)
    if rc != 0:
        raise ValueError(f"oracle rejected input (code {rc})")
    return out


def ref_cert_py(family: int, prims: np.ndarray, eps_time=1.0e-5, eps_space=1.0e-8, max_depth=32) -> np.ndarray:
    """Restated Python twin (oracle/ref_cert.c, pure_exact_cpu.py:98-308) over
    [0, 1]; `nodes` = exact evals / 3, `reason_code` = deepest depth."""
    prims = np.ascontiguousarray(prims, dtype=PRIM)
    out = np.zeros(prims.shape[0], RESULT)
    rc = lib().orc_ref_eval_py_batch(family, prims.ctypes.data, prims.shape[0], eps_time, eps_space, max_depth,
                                     out.ctypes.data)
    if rc != 0:
        raise ValueError(f"oracle rejected input (code {rc})")
    return out


def ref_cert_compiled(family: int, prims: np.ndarray, t0=0.0, t1=1.0, eps_time=1.0e-4, eps_space=1.0e-6,
                      max_depth=32) -> np.ndarray:
    """The reference's own compiled certificate_engine.cpp (oracle/_ref)."""
    prims = np.ascontiguousarray(prims, dtype=PRIM)
    out = np.zeros(prims.shape[0], RESULT)
    rc = ref_lib().ref_eval_batch(family, prims.ctypes.data, prims.shape[0], t0, t1, eps_time, eps_space,
                                  max_depth, out.ctypes.data)
    if rc != 0:
        raise ValueError(f"reference rejected primitive {rc - 1}")
    return out


def ref_engine_compiled(mask: int, pts: np.ndarray, ees: np.ndarray, t0=0.0, t1=1.0, eps_time=1.0e-4,
                        eps_space=1.0e-6, max_depth=32) -> np.ndarray | None:
    pts = np.ascontiguousarray(pts, dtype=PRIM)
    ees = np.ascontiguousarray(ees, dtype=PRIM)
    out = np.zeros(9, np.float64)
    rc = ref_lib().ref_engine_eval(mask, t0, t1, pts.ctypes.data, pts.shape[0], ees.ctypes.data, ees.shape[0],
                                   eps_time, eps_space, max_depth, out.ctypes.data)
    return None if rc != 0 else out


def tuv_cert(is_ee: bool, x: np.ndarray, ms=0.0, tolerance=1e-6, max_itr=1_000_000, heap_cap=1 << 20,
             threads=1) -> np.ndarray:
    """Canonical (t,u,v) inclusion certificate over (n, 8, 3) queries (oracle/tuv_cert.c)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(x.shape[0], TUV_RESULT)
    rc = lib().orc_tuv_eval_batch(int(is_ee), x.ctypes.data, x.shape[0], ms, tolerance, 1.0, max_itr, heap_cap,
                                  out.ctypes.data, threads)
    if rc != 0:
        raise ValueError(f"oracle rejected input (code {rc})")
    return out


def tuv_params(is_ee: bool, x: np.ndarray, tolerance=1e-6):
    err, inv = np.zeros(3), np.zeros(3)
    lib().orc_tuv_params(int(is_ee), np.ascontiguousarray(x, np.float64).ctypes.data, tolerance, err.ctypes.data,
                         inv.ctypes.data)
    return err, inv


# ---- broad phase ----------------------------------------------------------
def _bp_sigs():
    L = lib()
    if getattr(L, "_bp_ready", False):
        return L
    L.orc_swept_boxes.argtypes = [_vp, _vp, _i64, _vp, _i64, _vp, _i64, _f64, _vp, _vp, _vp]
    L.orc_swept_boxes.restype = None
    for name in ("orc_sweep_vf", "orc_brute_vf"):
        fn = getattr(L, name)
        fn.argtypes = [_vp, _vp, _i64, _vp, _vp, _vp, _i64, _vp, _i32, _vp, _i64]
        fn.restype = _i64
    for name in ("orc_sweep_ee", "orc_brute_ee"):
        fn = getattr(L, name)
        fn.argtypes = [_vp, _vp, _vp, _i64, _vp, _i32, _vp, _i64]
        fn.restype = _i64
    L._bp_ready = True
    return L


def swept_boxes(x0, x1, faces, edges, thickness):
    L = _bp_sigs()
    x0, x1 = np.ascontiguousarray(x0, np.float64), np.ascontiguousarray(x1, np.float64)
    faces, edges = np.ascontiguousarray(faces, np.int32), np.ascontiguousarray(edges, np.int32)
    vb, fb, eb = np.zeros((len(x0), 6)), np.zeros((len(faces), 6)), np.zeros((len(edges), 6))
    L.orc_swept_boxes(x0.ctypes.data, x1.ctypes.data, len(x0), faces.ctypes.data, len(faces), edges.ctypes.data,
                      len(edges), thickness, vb.ctypes.data, fb.ctypes.data, eb.ctypes.data)
    return vb, fb, eb


def object_pair_allowed(vbox, vobj, nobj):
    """Object swept boxes and the allowed matrix (PrepareScene :528-550):
    self pairs allowed, distinct objects when their boxes overlap."""
    ob = np.zeros((nobj, 6))
    ob[:, :3], ob[:, 3:] = np.inf, -np.inf
    for o in range(nobj):
        m = vobj == o
        if m.any():
            ob[o, :3] = vbox[m, :3].min(0)
            ob[o, 3:] = vbox[m, 3:].max(0)
    allowed = np.zeros((nobj, nobj), np.uint8)
    for a in range(nobj):
        for b in range(a, nobj):
            ok = a == b or bool(np.all(ob[a, :3] <= ob[b, 3:]) and np.all(ob[b, :3] <= ob[a, 3:]))
            allowed[a, b] = allowed[b, a] = ok
    return ob, allowed


def candidates(scene, method="sweep", cap=None):
    """VF and EE candidate pairs of a scene (sorted lexicographically)."""
    L = _bp_sigs()
    t = scene.topo
    vb, fb, eb = swept_boxes(scene.x0, scene.x1, t.faces, t.edges, scene.thickness)
    # the object box of PrepareScene is built from raw positions then expanded
    # by the thickness; it equals the union of the expanded vertex boxes
    _, allowed = object_pair_allowed(vb, t.vertex_object, t.object_count)
    vobj, fobj, eobj = (np.ascontiguousarray(a, np.int32) for a in (t.vertex_object, t.face_object, t.edge_object))
    out = {}
    for kind in ("vf", "ee"):
        c = cap or 1
        while True:
            buf = np.zeros((c, 2), np.int32)
            if kind == "vf":
                fn = L.orc_sweep_vf if method == "sweep" else L.orc_brute_vf
                n = fn(vb.ctypes.data, vobj.ctypes.data, len(vb), fb.ctypes.data, t.faces.ctypes.data, fobj.ctypes.data,
                       len(fb), allowed.ctypes.data, t.object_count, buf.ctypes.data, c)
            else:
                fn = L.orc_sweep_ee if method == "sweep" else L.orc_brute_ee
                n = fn(eb.ctypes.data, t.edges.ctypes.data, eobj.ctypes.data, len(eb), allowed.ctypes.data,
                       t.object_count, buf.ctypes.data, c)
            if n <= c:
                break
            c = int(n)
        pairs = buf[:n]
        out[kind] = pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))]
    return out


# ---- TinySTPF scorer --------------------------------------------------------
def tiny_features(is_vf: bool, rows: np.ndarray, index: np.ndarray) -> np.ndarray:
    """32 features per candidate from rows (n, 8, 3) and candidate indices (oracle/mlp.c)."""
    rows = np.ascontiguousarray(rows, np.float64)
    out = np.zeros((rows.shape[0], 32))
    for c in range(rows.shape[0]):
        f = np.zeros(32)
        lib().orc_tiny_features(int(is_vf), rows[c].ctypes.data, int(index[c]), f.ctypes.data)
        out[c] = f
    return out


def tiny_predict(weights: np.ndarray, feats: np.ndarray) -> np.ndarray:
    w = np.ascontiguousarray(weights, np.float64)
    feats = np.ascontiguousarray(feats, np.float64)
    return np.array([lib().orc_tiny_predict(w.ctypes.data, feats[c].ctypes.data) for c in range(feats.shape[0])])


def tiny_scores(policy: int, is_vf: bool, rows: np.ndarray, index: np.ndarray, keys: np.ndarray,
                weights: np.ndarray | None) -> np.ndarray:
    rows = np.ascontiguousarray(rows, np.float64)
    index = np.ascontiguousarray(index, np.uint64)
    keys = np.ascontiguousarray(keys, np.uint64)
    w = np.ascontiguousarray(weights if weights is not None else np.zeros(4289), np.float64)
    out = np.zeros(rows.shape[0])
    lib().orc_tiny_score_batch(policy, int(is_vf), rows.ctypes.data, index.ctypes.data, keys.ctypes.data,
                               rows.shape[0], w.ctypes.data, out.ctypes.data)
    return out


def set_mlp_mutation(k: int) -> None:
    ctypes.c_int.in_dll(lib(), "orc_mlp_mutation").value = k


# ---- proxy-scene candidates ---------------------------------------------------
RAW_HIT = np.dtype([("query_id", "<u8"), ("pair_key", "<u8"), ("proxy_a_index", "<u4"), ("proxy_b_index", "<u4"),
                    ("slab_id", "<u4"), ("rt_hit_count", "<u4"), ("flags", "<u4"), ("reserved0", "<u4"),
                    ("motion_bound", "<f4", (4,)), ("reserved1", "<u4"), ("reserved2", "<u4")])
CANDIDATE = np.dtype([("candidate_id", "<u8"), ("query_id", "<u8"), ("slab_id", "<u4"), ("object_a_id", "<u4"),
                      ("patch_a_id", "<u4"), ("object_b_id", "<u4"), ("patch_b_id", "<u4"), ("proxy_type_a", "<u4"),
                      ("proxy_type_b", "<u4"), ("rt_hit_count", "<u4"), ("motion_bound", "<f4", (4,)),
                      ("flags", "<u4"), ("reserved", "<u4")])
DENSITY = np.dtype([("proxy_count", "<u8"), ("object_count", "<u8"), ("slab_count", "<u8"),
                    ("cross_object_same_slab_pair_count", "<u8"), ("raw_hit_count", "<u8"),
                    ("compact_candidate_count", "<u8"), ("raw_hits_per_proxy", "<f8"), ("candidates_per_proxy", "<f8"),
                    ("candidates_per_slab", "<f8"), ("aabb_overlap_ratio", "<f8"),
                    ("avg_rt_hits_per_candidate", "<f8")])
assert RAW_HIT.itemsize == 64 and CANDIDATE.itemsize == 72


class _ProxySet(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in ("object_id", "patch_id", "slab_id", "proxy_type", "t0", "t1",
                                                      "bounds", "motion")] + [("n", ctypes.c_int64)]


PROXY_FIELDS = ("object_id", "patch_id", "slab_id", "proxy_type", "t0", "t1", "bounds", "motion")


def _proxy_arrays(scene: dict):
    arr = {k: np.ascontiguousarray(scene[k], dtype=np.uint32) for k in PROXY_FIELDS[:4]}
    arr.update({k: np.ascontiguousarray(scene[k], dtype=np.float64) for k in PROXY_FIELDS[4:]})
    s = _ProxySet(*[arr[k].ctypes.data for k in PROXY_FIELDS], len(arr["t0"]))
    return s, arr


def _proxy_run(fn, query_id: int, scene: dict, with_err: bool):
    s, keep = _proxy_arrays(scene)
    n = len(keep["t0"])
    cap = max(n * (n - 1) // 2, 1)
    raw = np.zeros(cap, RAW_HIT)
    cand = np.zeros(cap, CANDIDATE)
    dens = np.zeros(1, DENSITY)
    nr, nc = ctypes.c_int64(), ctypes.c_int64()
    vp = ctypes.c_void_p
    args = [ctypes.c_uint64(query_id), ctypes.byref(s), ctypes.c_int64(cap), vp(raw.ctypes.data), ctypes.byref(nr),
            ctypes.c_int64(cap), vp(cand.ctypes.data), ctypes.byref(nc), vp(dens.ctypes.data)]
    if with_err:
        err = ctypes.create_string_buffer(256)
        rc = fn(*args, err, 256)
        if rc != 0:
            raise ValueError(err.value.decode())
    else:
        rc = fn(*args)
        if rc != 0:
            raise ValueError(f"oracle rejected the proxy scene (code {rc})")
    return raw[: nr.value], cand[: nc.value], dens[0]


def proxy_generate(query_id: int, scene: dict):
    """Raw hits, compact candidates and density of a proxy scene given as arrays (oracle/proxy.c)."""
    return _proxy_run(lib().orc_proxy_generate, query_id, scene, False)


def ref_proxy_generate(query_id: int, scene: dict):
    """The same through the reference's compiled candidate_buffer.cpp (oracle/_ref)."""
    return _proxy_run(ref_lib().ref_proxy_generate, query_id, scene, True)


def ref_proxy_kat_scene() -> tuple[int, dict]:
    L = ref_lib()
    cap = 16
    out = {k: np.zeros(cap, np.uint32) for k in PROXY_FIELDS[:4]}
    out.update(t0=np.zeros(cap), t1=np.zeros(cap), bounds=np.zeros((cap, 6)), motion=np.zeros((cap, 4)))
    n, q = ctypes.c_int64(), ctypes.c_uint64()
    rc = L.ref_proxy_kat_scene(ctypes.c_int64(cap), *[ctypes.c_void_p(out[k].ctypes.data) for k in PROXY_FIELDS],
                               ctypes.byref(n), ctypes.byref(q))
    if rc != 0:
        raise ValueError("reference BuildProxyScene failed")
    return int(q.value), {k: v[: n.value] for k, v in out.items()}


def random_proxy_scene(rng, n: int, n_objects: int, n_slabs: int, extent: float = 1.0, size: float = 0.15) -> dict:
    """A seeded proxy scene: boxes in [-extent, extent]^3, objects, patches, and
    slabs that partition [0, 1] (each slab one global interval)."""
    lo = rng.uniform(-extent, extent, (n, 3))
    hi = lo + rng.uniform(0.0, size, (n, 3))
    slab = rng.integers(0, n_slabs, n).astype(np.uint32)
    edges = np.linspace(0.0, 1.0, n_slabs + 1)
    return dict(object_id=rng.integers(1, n_objects + 1, n).astype(np.uint32),
                patch_id=rng.integers(0, 6, n).astype(np.uint32), slab_id=slab,
                proxy_type=rng.integers(1, 3, n).astype(np.uint32), t0=edges[slab], t1=edges[slab + 1],
                bounds=np.concatenate([lo, hi], 1), motion=rng.uniform(0.0, 0.1, (n, 4)))


# ---- proposal policy (compiled reference only) --------------------------------
def ref_schedule_runtime(cands, rows, outs, masks: dict, cfg: dict):
    """ScheduleRuntimeExactWorkItemsFromProposals of the compiled reference.
    Returns (u64[n, 8], f64[n, 2], f32[n], stats u64[10]) or raises ValueError."""
    L = ref_lib()
    vp = ctypes.c_void_p
    n = len(cands)
    c = lambda a, t: np.ascontiguousarray(a, dtype=t)  # noqa: E731
    keep = [c(cands["candidate_id"], np.uint64), c(cands["query_id"], np.uint64), c(cands["slab_id"], np.uint32),
            c(cands["object_a_id"], np.uint32), c(cands["patch_a_id"], np.uint32), c(cands["object_b_id"], np.uint32),
            c(cands["patch_b_id"], np.uint32), c(cands["proxy_type_a"], np.uint32), c(cands["proxy_type_b"], np.uint32),
            c(cands["rt_hit_count"], np.uint32), c(cands["motion_bound"], np.float32)]
    rk = [c(rows["candidate_id"], np.uint64), c(rows["features"], np.float32)]
    ok = [c(outs["candidate_id"], np.uint64), c(outs["interval_scores"], np.float32),
          c(outs["family_scores"], np.float32), c(outs["priority_score"], np.float32),
          c(outs["cost_score"], np.float32), c(outs["uncertainty_score"], np.float32)]
    mq = c(list(masks.keys()), np.uint64)
    mv = c(list(masks.values()), np.uint32)
    cd = c([cfg["fallback_interval_t0"], cfg["fallback_interval_t1"]], np.float64)
    cf = c([cfg["family_score_threshold"], cfg["uncertainty_fallback_threshold"], cfg["ood_abs_feature_threshold"]],
           np.float32)
    cu = c([cfg["first_work_item_id"], cfg["conservative_feature_family_mask"], int(cfg["preserve_candidate_order"])],
           np.uint64)
    wu = np.zeros((max(n, 1), 8), np.uint64)
    wi = np.zeros((max(n, 1), 2))
    wp = np.zeros(max(n, 1), np.float32)
    st = np.zeros(10, np.uint64)
    err = ctypes.create_string_buffer(256)
    rc = L.ref_schedule_runtime(ctypes.c_int64(n), *[vp(a.ctypes.data) for a in keep], ctypes.c_int64(len(rows)),
                                *[vp(a.ctypes.data) for a in rk], ctypes.c_int64(len(outs)),
                                *[vp(a.ctypes.data) for a in ok], ctypes.c_int64(len(mq)), vp(mq.ctypes.data),
                                vp(mv.ctypes.data), vp(cd.ctypes.data), vp(cf.ctypes.data), vp(cu.ctypes.data),
                                vp(wu.ctypes.data), vp(wi.ctypes.data), vp(wp.ctypes.data), vp(st.ctypes.data), err,
                                256)
    if rc != 0:
        raise ValueError(err.value.decode())
    return wu[:n], wi[:n], wp[:n], st


def ref_dense_group(inputs, cfg: dict):
    L = ref_lib()
    vp = ctypes.c_void_p
    n = len(inputs)
    c = lambda a, t: np.ascontiguousarray(a, dtype=t)  # noqa: E731
    d6 = c(np.stack([inputs[f] for f in ("full_exact_cost", "narrow_exact_cost", "interval_t0", "interval_t1",
                                         "contact_t0", "contact_t1")], 1), np.float64)
    f3 = c(np.stack([inputs[f] for f in ("priority_score", "cost_score", "uncertainty_score")], 1), np.float32)
    q, ci, col = c(inputs["query_id"], np.uint64), c(inputs["candidate_id"], np.uint64), \
        c(inputs["candidate_collides"], np.uint8)
    cf = c([cfg["uncertainty_fallback_threshold"]], np.float32)
    cu = c([cfg["representative_attempt_limit"], int(cfg["preserve_input_order"])], np.uint64)
    su, sd = np.zeros(14, np.uint64), np.zeros(4)
    err = ctypes.create_string_buffer(256)
    rc = L.ref_dense_group(ctypes.c_int64(n), vp(q.ctypes.data), vp(ci.ctypes.data), vp(d6.ctypes.data),
                           vp(f3.ctypes.data), vp(col.ctypes.data), vp(cf.ctypes.data), vp(cu.ctypes.data),
                           ctypes.c_double(cfg["interval_miss_penalty_scale"]), vp(su.ctypes.data), vp(sd.ctypes.data),
                           err, 256)
    if rc != 0:
        raise ValueError(err.value.decode())
    return su, sd


# ---- external query batches and proxy-scene construction ---------------------
EXT_QUERY = np.dtype({"names": ["source_query_id", "source_query_index", "family", "v0", "v1", "has_ground_truth",
                                "ground_truth_collides", "has_box_pair", "box_pair"],
                      "formats": ["<u8", "<u4", "<u4", ("<f8", (4, 3)), ("<f8", (4, 3)), "u1", "u1", "u1",
                                  ("<u4", (2,))],
                      "offsets": [0, 8, 12, 16, 112, 208, 209, 210, 212], "itemsize": 224})
EXT_STATS = np.dtype([("primitive_count", "<u8"), ("raw_hit_count", "<u8"), ("compact_candidate_count", "<u8"),
                      ("candidate_recall", "<f8")])
BUILD_OBJECT = np.dtype({"names": ["object_id", "proxy_type", "slabs_per_motion_segment", "n_patches", "n_segments",
                                   "patch0", "segment0", "eps_proxy"],
                         "formats": ["<u4"] * 7 + ["<f8"], "offsets": [0, 4, 8, 12, 16, 20, 24, 32], "itemsize": 40})
BUILD_PATCH = np.dtype({"names": ["patch_id", "local_center", "radius"], "formats": ["<u4", ("<f8", (3,)), "<f8"],
                        "offsets": [0, 8, 32], "itemsize": 40})
BUILD_SEGMENT = np.dtype([("t0", "<f8"), ("t1", "<f8"), ("translation0", "<f8", (3,)), ("rotation0", "<f8", (4,)),
                          ("translation1", "<f8", (3,)), ("rotation1", "<f8", (4,))])
PRIM_DOUBLES = 36  # one built proxy primitive, see ref_external_shim.cpp


def ref_external_batch(queries: np.ndarray):
    """GenerateCandidatesForExternalBatch of the compiled reference:
    (candidates CANDIDATE[k], runtime ids u64[n], stats EXT_STATS) or ValueError."""
    L = ref_lib()
    q = np.ascontiguousarray(queries, dtype=EXT_QUERY)
    n = len(q)
    cand = np.zeros(max(n, 1), CANDIDATE)
    rid = np.zeros(max(n, 1), np.uint64)
    st = np.zeros(1, EXT_STATS)
    nc = ctypes.c_int64()
    err = ctypes.create_string_buffer(256)
    vp = ctypes.c_void_p
    rc = L.ref_external_batch(ctypes.c_int64(n), vp(q.ctypes.data), ctypes.c_int64(len(cand)), vp(cand.ctypes.data),
                              ctypes.byref(nc), vp(rid.ctypes.data), vp(st.ctypes.data), err, 256)
    if rc != 0:
        raise ValueError(err.value.decode())
    return cand[: nc.value], rid[:n], st[0]


def ref_build_proxy_scene(query_id: int, objects: np.ndarray, patches: np.ndarray, segments: np.ndarray):
    """BuildProxyScene of the compiled reference: float64[n, 36] (the field
    order of ref_external_shim.cpp) or ValueError."""
    L = ref_lib()
    o = np.ascontiguousarray(objects, dtype=BUILD_OBJECT)
    p = np.ascontiguousarray(patches, dtype=BUILD_PATCH)
    s = np.ascontiguousarray(segments, dtype=BUILD_SEGMENT)
    cap = 4096
    out = np.zeros((cap, PRIM_DOUBLES))
    n = ctypes.c_int64()
    err = ctypes.create_string_buffer(256)
    vp = ctypes.c_void_p
    rc = L.ref_build_proxy_scene(ctypes.c_uint64(query_id), ctypes.c_int(len(o)), vp(o.ctypes.data),
                                 vp(p.ctypes.data), vp(s.ctypes.data), ctypes.c_int64(cap), vp(out.ctypes.data),
                                 ctypes.byref(n), err, 256)
    if rc != 0:
        raise ValueError(err.value.decode())
    return out[: n.value]


def external_batch(queries: np.ndarray):
    """Restatement of GenerateCandidatesForExternalBatch
    (external_batch_candidate.cpp:128-339) taken literally: the batched proxy
    scene (two swept-AABB proxies per query, one slab [2q, 2q + 1] each,
    shifted along x by the running extent), every proxy pair i < j tested as
    GenerateRawCandidatesCpu does (candidate_buffer.cpp:144-181), then the
    translation, the sort by (query_id, patch_a_id, patch_b_id) and
    MakeCandidateId.  O(n^2): small batches only.  Returns (candidates as a
    list of dicts, runtime ids)."""
    with np.errstate(over="ignore"):  # overflow is one of the reference's error cases
        return _external_batch(np.asarray(queries, dtype=EXT_QUERY))


def _external_batch(q: np.ndarray):
    n = len(q)
    seen, taken, rids = set(), set(), []
    cursor = 0.0
    prox = []  # (query index, object, slab, t0, t1, lo[3], hi[3])
    for i in range(n):
        if not np.isfinite(q["v0"][i]).all():
            raise ValueError("external batch query vertices_t0 must be finite")
        if not np.isfinite(q["v1"][i]).all():
            raise ValueError("external batch query vertices_t1 must be finite")
        sid = int(q["source_query_id"][i])
        if sid in seen:
            raise ValueError("external batch source_query_id values must be unique")
        seen.add(sid)
        rid = sid if sid > 0 else i + 1
        while rid in taken:
            rid += 1
        taken.add(rid)
        rids.append(rid)
        sides = ([0], [1, 2, 3]) if q["family"][i] == 0 else ([0, 1], [2, 3])
        boxes = []
        for idx in sides:
            pts = np.concatenate([q["v0"][i][idx], q["v1"][i][idx]])
            boxes.append((pts.min(0), pts.max(0)))
        lo_x = min(boxes[0][0][0], boxes[1][0][0])
        hi_x = max(boxes[0][1][0], boxes[1][1][0])
        off = cursor - lo_x
        t0 = float(i) * 2.0
        for obj, (lo, hi) in zip((1, 2), boxes):
            lo, hi = lo.copy(), hi.copy()
            lo[0] += off
            hi[0] += off
            if not (np.isfinite(lo).all() and np.isfinite(hi).all()):
                raise ValueError("proxy bounds contain a non-finite value")
            prox.append((i, obj, i, t0, t0 + 1.0, lo, hi))
        cursor += (hi_x - lo_x) + 1.0e-2
    hits = []
    for a in range(len(prox)):
        for b in range(a + 1, len(prox)):
            A, B = prox[a], prox[b]
            if A[1] == B[1] or A[2] != B[2] or not (A[3] < B[4] and B[3] < A[4]):
                continue
            if (A[6] < B[5]).any() or (B[6] < A[5]).any():
                continue
            rad = [0.5 * max(P[6] - P[5]) for P in (A, B)]
            r = np.float32(max(rad))
            if not np.isfinite(r) or r < 0:
                raise ValueError("raw candidate motion_bound must be finite and non-negative")
            hits.append((A[0], r))
    out = []
    for i, r in hits:
        out.append(dict(query_id=rids[i], slab_id=i,
                        object_a_id=int(q["box_pair"][i][0]) if q["has_box_pair"][i] else 1,
                        patch_a_id=0 if q["family"][i] == 0 else 1,
                        object_b_id=int(q["box_pair"][i][1]) if q["has_box_pair"][i] else 2,
                        patch_b_id=(int(q["source_query_index"][i]) + 1) & 0xFFFFFFFF, motion_bound=r))
    out.sort(key=lambda c: (c["query_id"], c["patch_a_id"], c["patch_b_id"]))
    for k, c in enumerate(out):
        c["candidate_id"] = (c["query_id"] * 1_000_003 + k + 1) & 0xFFFFFFFFFFFFFFFF
    return out, np.array(rids, dtype=np.uint64)


# ---- exact work queue (oracle/ref_queue.c, ref_shim.cpp) ---------------------
# Byte layouts of the library's p2c_queue_entry / p2c_certificate / p2c_audit_row.
QUEUE_ENTRY = np.dtype([
    ("work_item_id", "<u8"), ("parent_candidate_id", "<u8"), ("query_id", "<u8"), ("slab_id", "<u4"),
    ("patch_a_id", "<u4"), ("patch_b_id", "<u4"), ("feature_family_mask", "<u4"), ("interval_t0", "<f8"),
    ("interval_t1", "<f8"), ("topk_feature_ids_offset", "<u4"), ("depth", "<u2"), ("source", "u1"),
    ("pad0", "u1"), ("priority_score", "<f4"), ("max_subdivision_depth", "<u4"), ("eps_time", "<f8"),
    ("eps_space", "<f8"), ("pt_begin", "<i8"), ("pt_count", "<i8"), ("ee_begin", "<i8"), ("ee_count", "<i8"),
])
CERTIFICATE = np.dtype([
    ("work_item_id", "<u8"), ("query_id", "<u8"), ("interval_t0", "<f8"), ("interval_t1", "<f8"),
    ("toi_upper", "<f8"), ("safe_margin_lb", "<f8"), ("witness_id_a", "<i8"), ("witness_id_b", "<i8"),
    ("eps_time", "<f8"), ("eps_space", "<f8"), ("covered_feature_mask", "<u4"), ("reason_code", "<u2"),
    ("status", "u1"), ("witness_family", "u1"), ("next_refinement_mode", "u1"), ("pad", "u1", (7,)),
])
AUDIT_ROW = np.dtype([
    ("event_id", "<u8"), ("query_id", "<u8"), ("candidate_id", "<u8"), ("work_item_id", "<u8"),
    ("interval_t0", "<f8"), ("interval_t1", "<f8"), ("timestamp_us", "<u8"), ("aux_value0", "<f8"),
    ("aux_value1", "<f8"), ("action", "<u2"), ("depth", "<u2"), ("stage", "u1"), ("pad", "u1", (3,)),
])
assert QUEUE_ENTRY.itemsize == 120 and CERTIFICATE.itemsize == 96 and AUDIT_ROW.itemsize == 80
_u64 = ctypes.c_uint64


def _queue_sigs(L, compiled: bool):
    if getattr(L, "_queue_ready", False):
        return L
    if compiled:
        L.ref_process_queue.argtypes = [_vp, _i64, _vp, _vp, _u64, _u64, _i32, _vp, _vp, ctypes.c_char_p, _i64]
        L.ref_process_queue.restype = _i32
        L.ref_refine_queue.argtypes = [_vp, _vp, _i64, _u64, ctypes.c_uint32, _f64, _vp, _i64, ctypes.c_char_p,
                                       _i64]
        L.ref_refine_queue.restype = _i64
    else:
        L.orc_queue_process.argtypes = [_vp, _i64, _vp, _vp, _u64, _u64, _i32, _vp, _vp,
                                        ctypes.POINTER(ctypes.c_char_p)]
        L.orc_queue_process.restype = _i32
        L.orc_queue_refine.argtypes = [_vp, _vp, _i64, _u64, ctypes.c_uint32, _f64, _vp, _i64]
        L.orc_queue_refine.restype = _i64
    L._queue_ready = True
    return L


def _queue_args(entries, pts, ees):
    entries = np.ascontiguousarray(entries, QUEUE_ENTRY)
    pts = np.ascontiguousarray(pts, PRIM)
    ees = np.ascontiguousarray(ees, PRIM)
    return entries, pts, ees


def queue_process(entries, pts, ees, first_event_id=1, first_timestamp_us=1, emit=True, compiled=False):
    """ProcessExactWorkQueueCpu: (certificates, rows), or ValueError with the
    reference's message.  compiled=True runs the reference's own code
    (oracle/_ref), else the C restatement."""
    entries, pts, ees = _queue_args(entries, pts, ees)
    n = entries.shape[0]
    certs = np.zeros(n, CERTIFICATE)
    rows = np.zeros(n * (2 if emit else 1), AUDIT_ROW)
    if compiled:
        L = _queue_sigs(ref_lib(), True)
        msg = ctypes.create_string_buffer(512)
        rc = L.ref_process_queue(entries.ctypes.data, n, pts.ctypes.data, ees.ctypes.data, first_event_id,
                                 first_timestamp_us, int(emit), certs.ctypes.data, rows.ctypes.data, msg, 512)
        text = msg.value.decode()
    else:
        L = _queue_sigs(lib(), False)
        m = ctypes.c_char_p()
        rc = L.orc_queue_process(entries.ctypes.data, n, pts.ctypes.data, ees.ctypes.data, first_event_id,
                                 first_timestamp_us, int(emit), certs.ctypes.data, rows.ctypes.data, ctypes.byref(m))
        text = (m.value or b"").decode()
    if rc != 0:
        raise ValueError(text)
    return certs, rows


def queue_refine(entries, certs, first_child=1, max_child_depth=64, min_interval_width=1.0e-4, compiled=False):
    """GenerateConservativeRefinementWorkItems per entry in queue order with a
    running first child id; children as entries."""
    entries = np.ascontiguousarray(entries, QUEUE_ENTRY)
    certs = np.ascontiguousarray(certs, CERTIFICATE)
    n = entries.shape[0]
    out = np.zeros(2 * n, QUEUE_ENTRY)
    if compiled:
        L = _queue_sigs(ref_lib(), True)
        msg = ctypes.create_string_buffer(512)
        got = L.ref_refine_queue(entries.ctypes.data, certs.ctypes.data, n, first_child, max_child_depth,
                                 min_interval_width, out.ctypes.data, 2 * n, msg, 512)
        if got < 0:
            raise ValueError(msg.value.decode())
    else:
        L = _queue_sigs(lib(), False)
        got = L.orc_queue_refine(entries.ctypes.data, certs.ctypes.data, n, first_child, max_child_depth,
                                 min_interval_width, out.ctypes.data, 2 * n)
    return out[:got]
