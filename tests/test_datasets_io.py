"""On-disk inputs: the Tight-Inclusion rational CSV (the reference package's
reader, src/python/p2cccd/datasets/tight_inclusion_queries.py:14-290, and the
C++ benchmark's rule, tight_inclusion_full_query_benchmark.cpp:291-321,
412-455) and binary PLY frames (scene_object_envelope_native_ti_walltime.cpp:
338-489), on hand-written fixtures.  Expected coordinates come from exact
rational arithmetic with an independent round-to-nearest-even helper: the
Python rule rounds numerator and denominator to 53 bits and the quotient to
53; the C++ rule rounds them to 64 bits (x87 long double), the quotient to 64
and then to 53."""
import struct
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

from p2cccd import datasets
from p2cccd.datasets import tight_inclusion_queries as ti

cpp = pytest.importorskip("p2cccd_cpp")
FIX = Path(__file__).resolve().parent / "golden" / "tight_inclusion"
VF = FIX / "handwritten" / "vertex-face" / "queries.csv"
EE = FIX / "handwritten" / "edge-edge" / "queries.csv"


def round_bits(q: Fraction, p: int) -> Fraction:
    """q rounded to a p-bit significand, ties to even (no exponent limits)."""
    if q == 0:
        return Fraction(0)
    s, q = (-1 if q < 0 else 1), abs(q)
    e = q.numerator.bit_length() - q.denominator.bit_length()
    while Fraction(2) ** e > q:
        e -= 1
    while Fraction(2) ** (e + 1) <= q:
        e += 1
    m = q * Fraction(2) ** (p - 1 - e)
    f = m.numerator // m.denominator
    r = m - f
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and f % 2):
        f += 1
    return s * Fraction(f) / Fraction(2) ** (p - 1 - e)


def expected(path: Path, rule: str) -> np.ndarray:
    rows = [l.strip().split(",") for l in path.read_text().splitlines()]
    out = []
    for r in rows:
        xyz = []
        for k in range(3):
            n, d = Fraction(int(r[2 * k])), Fraction(int(r[2 * k + 1]))
            if rule == "python":
                xyz.append(float(round_bits(round_bits(n, 53) / round_bits(d, 53), 53)))
            else:
                xyz.append(float(round_bits(round_bits(round_bits(n, 64) / round_bits(d, 64), 64), 53)))
        out.append(xyz)
    return np.array(out).reshape(-1, 8, 3)


@pytest.mark.parametrize("path,truth", [(VF, [1, 0, 0]), (EE, [1, 0])])
@pytest.mark.parametrize("rule", ["python", "cpp"])
def test_tight_inclusion_fixture_values(path, truth, rule):
    x, t = ti.load_queries(path, rule=rule)
    assert x.shape == (len(truth), 8, 3) and list(t) == [bool(v) for v in truth]
    assert np.array_equal(x, expected(path, rule))


def test_rules_differ_where_rounding_does():
    xp, _ = ti.load_queries(VF, rule="python")
    xc, _ = ti.load_queries(VF, rule="cpp")
    # (2^53 + 1) / (2^55 + 12): both operands need more than 53 bits
    assert xp[2, 0, 0] != xc[2, 0, 0]
    same = np.ones(xp.shape, bool)
    same[2, 0, 0] = same[2, 4, 0] = False
    assert np.array_equal(xp[same], xc[same])


def test_reference_api_objects():
    files = datasets.discover_tight_inclusion_csv_files(FIX, inspect=True)
    assert [(f.case_name, f.kind, f.query_count, f.line_count) for f in files] == [
        ("handwritten", "vertex-face", 3, 24), ("handwritten", "edge-edge", 2, 16)]
    q = datasets.read_tight_inclusion_query(VF, 2, dataset_root=FIX)
    assert q.case_name == "handwritten" and q.kind == "vertex-face" and q.query_index == 2 and not q.ground_truth
    assert q.numerators[0, 0] == 2**53 + 1 and q.denominators[0, 0] == 2**55 + 12
    assert q.numerators[1, 2] == -2 and q.denominators[1, 2] == -7
    assert np.array_equal(q.vertices_t0, q.vertices_t0_t1[:4]) and np.array_equal(q.vertices_t1, q.vertices_t0_t1[4:])
    f = q.rational_magnitude_features
    assert f[0] == pytest.approx(np.log1p(float(2**53 + 1))) and f[1] == pytest.approx(np.log1p(float(2**55 + 12)))
    qs = list(datasets.iter_tight_inclusion_queries(VF, start_query_index=1, limit=2))
    assert [x.query_index for x in qs] == [1, 2]
    big = ti._log1p_abs(10**400)
    assert big == pytest.approx(401 * np.log(10.0))


def _write(tmp_path, name, text, kind="vertex-face"):
    d = tmp_path / "case" / kind
    d.mkdir(parents=True, exist_ok=True)
    p = d / name
    p.write_text(text)
    return p


GOOD = "1,1,0,1,0,1,0\n"


@pytest.mark.parametrize("row,msg", [
    ("1,1,0,1,0,1\n", "expected 7 columns at line 1, got 6"),
    ("1,1,0,1,0,x,0\n", "non-integer rational CSV value at line 1: 1,1,0,1,0,x,0"),
    ("1,0,0,1,0,1,0\n", "zero denominator in Tight-Inclusion CSV row at line 1"),
    ("1,1,0,1,0,1,2\n", "ground-truth value must be 0 or 1 at line 1, got 2"),
])
def test_python_rule_row_errors(tmp_path, row, msg):
    p = _write(tmp_path, "bad.csv", row + GOOD * 7)
    with pytest.raises(ValueError) as e:
        datasets.read_tight_inclusion_query(p, 0)
    assert str(e.value) == msg


def test_python_rule_block_errors(tmp_path):
    p = _write(tmp_path, "mixed.csv", GOOD * 7 + "1,1,0,1,0,1,1\n")
    with pytest.raises(ValueError, match=r"inconsistent truth labels in .*mixed.csv:0"):
        datasets.read_tight_inclusion_query(p, 0)
    p = _write(tmp_path, "short.csv", GOOD * 11)
    with pytest.raises(ValueError, match="truncated query block with 3 leftover rows"):
        list(datasets.iter_tight_inclusion_queries(p))
    with pytest.raises(ValueError, match="line_count=11 is not divisible by 8"):
        datasets.inspect_tight_inclusion_csv(p)
    with pytest.raises(ValueError, match="query_index=1 is out of range"):
        datasets.read_tight_inclusion_query(p, 1)
    with pytest.raises(ValueError, match="unsupported Tight-Inclusion query kind"):
        datasets.normalize_query_kind("face-face")
    with pytest.raises(ValueError, match="cannot infer Tight-Inclusion query kind"):
        datasets.infer_query_kind(Path("/tmp/x/queries.csv"))


def test_cpp_rule_quirks_and_errors(tmp_path):
    # the benchmark's splitter keeps no field after a final ',' and trims '\r', ' ', '\t'
    p = _write(tmp_path, "q.csv", "1,2,3,4,5,8,0,\n" * 4 + "1,2,3,4,5,8,0 \t\r\n" * 4 + GOOD * 3)
    x, t = cpp.read_tight_inclusion_csv(str(p))
    assert x.shape == (1, 8, 3) and list(x[0, 0]) == [0.5, 0.75, 0.625] and not t[0]  # trailing block ignored
    # a '\r' after a final ',' is a field of its own (trimmed to empty): 8 fields
    p = _write(tmp_path, "cr.csv", "1,2,3,4,5,8,0,\r\n" * 8)
    with pytest.raises(RuntimeError, match="must contain 7 columns"):
        cpp.read_tight_inclusion_csv(str(p))
    p = _write(tmp_path, "six.csv", "1,1,0,1,0,1\n" * 8)
    with pytest.raises(RuntimeError, match="must contain 7 columns"):
        cpp.read_tight_inclusion_csv(str(p))
    p = _write(tmp_path, "truth.csv", "1,1,0,1,0,1,yes\n" * 8)
    with pytest.raises(RuntimeError, match="truth column must be 0 or 1"):
        cpp.read_tight_inclusion_csv(str(p))
    p = _write(tmp_path, "mixed.csv", GOOD * 7 + "1,1,0,1,0,1,1\n")
    with pytest.raises(RuntimeError, match="truth label changes inside query block"):
        cpp.read_tight_inclusion_csv(str(p))
    x, _ = cpp.read_tight_inclusion_csv(str(EE), 1)
    assert x.shape == (1, 8, 3)


# ---- binary PLY -------------------------------------------------------------
_FMT = {"float": "f", "double": "d", "uchar": "B", "char": "b", "ushort": "H", "short": "h", "uint": "I", "int": "i",
        "float32": "f", "float64": "d", "uint8": "B", "int32": "i", "uint32": "I", "uint16": "H"}


def write_ply(path, verts, faces, *, big=False, vprops=(("x", "float"), ("y", "float"), ("z", "float")),
              count_type="uchar", index_type="int", extra_face_element=False, fmt=None):
    e = ">" if big else "<"
    head = ["ply", f"format {fmt or ('binary_big_endian' if big else 'binary_little_endian')} 1.0",
            "comment hand-written fixture", f"element vertex {len(verts)}"]
    head += [f"property {t} {n}" for n, t in vprops]
    head += [f"element face {len(faces)}", f"property list {count_type} {index_type} vertex_indices"]
    if extra_face_element:
        head += ["element edge 0", "property int vertex1"]
    head.append("end_header")
    body = b""
    for v in verts:
        vals = {"x": v[0], "y": v[1], "z": v[2]}
        for n, t in vprops:
            body += struct.pack(e + _FMT[t], vals.get(n, 7))
    for f in faces:
        body += struct.pack(e + _FMT[count_type], len(f))
        for i in f:
            body += struct.pack(e + _FMT[index_type], i)
    path.write_bytes(("\n".join(head) + "\n").encode() + body)
    return path


QUAD = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (2, 0, 0), (3, 0, 0), (3, 1, 0), (2.5, 1.5, 0), (2, 1, 0)]
FACES = [(0, 1, 2, 3), (0, 0, 1), (4, 5, 6, 7, 8), (1, 2)]
FANS = [(0, 1, 2), (0, 2, 3), (4, 5, 6), (4, 6, 7), (4, 7, 8)]


@pytest.mark.parametrize("big", [False, True])
@pytest.mark.parametrize("vtype,ctype,itype", [("float", "uchar", "int"), ("double", "ushort", "uint"),
                                               ("int", "uint8", "int32")])
def test_ply_fans_and_types(tmp_path, big, vtype, ctype, itype):
    vprops = (("nx", "uchar"), ("x", vtype), ("y", vtype), ("z", vtype), ("red", "short"))
    verts = [(int(a * 2), int(b * 2), c) for a, b, c in QUAD] if vtype == "int" else QUAD
    p = write_ply(tmp_path / "m.ply", verts, FACES, big=big, vprops=vprops, count_type=ctype, index_type=itype,
                  extra_face_element=True)
    v, f = cpp.load_binary_ply(str(p))
    assert np.array_equal(v, np.array(verts, dtype=np.float64)) and f.dtype == np.int32
    assert [tuple(x) for x in f] == FANS


def test_ply_errors(tmp_path):
    p = tmp_path / "a.ply"
    p.write_bytes(b"plx\nend_header\n")
    with pytest.raises(RuntimeError, match="not a PLY file"):
        cpp.load_binary_ply(str(p))
    write_ply(p, QUAD, FACES, fmt="ascii")
    with pytest.raises(RuntimeError, match="only binary PLY is supported"):
        cpp.load_binary_ply(str(p))
    write_ply(p, QUAD, [])
    with pytest.raises(RuntimeError, match="PLY has no vertices or faces"):
        cpp.load_binary_ply(str(p))
    write_ply(p, QUAD, FACES)
    p.write_bytes(p.read_bytes()[:-3])
    with pytest.raises(RuntimeError, match="unexpected EOF while reading binary PLY"):
        cpp.load_binary_ply(str(p))
    q = write_ply(tmp_path / "b.ply", QUAD, FACES, vprops=(("x", "float"), ("y", "float"), ("z", "float")))
    q.write_bytes(q.read_bytes().replace(b"property float z", b"property half  z"))
    with pytest.raises(RuntimeError, match="unsupported PLY scalar type: half"):
        cpp.load_binary_ply(str(q))
    r = write_ply(tmp_path / "c.ply", QUAD, FACES)
    r.write_bytes(r.read_bytes().replace(b"list uchar int", b"list uchar f32"))
    with pytest.raises(RuntimeError, match="unsupported PLY integer type: f32"):
        cpp.load_binary_ply(str(r))


def test_scene_from_ply_frames_objects_and_edges(tmp_path):
    from p2cccd.datasets.ply import scene_from_ply_frames

    # three tetrahedra, one per object, written as t = 0 and t = 1 frames
    tet = [(0, 1, 2), (0, 3, 1), (1, 3, 2), (2, 3, 0)]
    base = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)], float)
    verts0 = np.concatenate([base + (3 * k, 0, 0) for k in (2, 0, 1)])
    faces = [tuple(int(i) + 4 * k for i in f) for k in range(3) for f in tet]
    f0 = write_ply(tmp_path / "f0.ply", verts0.tolist(), faces, vprops=(("x", "double"), ("y", "double"),
                                                                        ("z", "double")))
    f1 = write_ply(tmp_path / "f1.ply", (verts0 + 0.5).tolist(), faces, vprops=(("x", "double"), ("y", "double"),
                                                                               ("z", "double")))
    s = scene_from_ply_frames(f0, f1, thickness=0.01)
    assert s.topo.object_count == 3 and list(s.topo.vertex_object) == [0] * 4 + [1] * 4 + [2] * 4
    assert s.topo.edges.shape == (18, 2) and list(map(tuple, s.topo.edges[:6])) == [
        (0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]
    assert np.array_equal(s.x1 - s.x0, np.full_like(s.x0, 0.5))
    g = write_ply(tmp_path / "g.ply", verts0[:8].tolist(), faces[:8], vprops=(("x", "double"), ("y", "double"),
                                                                            ("z", "double")))
    with pytest.raises(ValueError, match="frame vertex counts differ"):
        scene_from_ply_frames(f0, g)
