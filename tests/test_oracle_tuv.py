"""Pins the (t,u,v) inclusion-certificate oracle (oracle/tuv_cert.c) without
the library it restates (Tight-Inclusion, not available here, so bit parity
with it stays unpinned):

  * closed forms: axis-aligned crossings whose contact time is z0 / (z0 - z1);
  * exact-rational brute force (sympy): the first root in [0,1] of the
    coplanarity cubic at which the primitives actually meet; a contact must
    be reported no later than it, and every reported time must find the
    primitives within the tolerance of touching;
  * the error bound: exact corner values of F (rational arithmetic) lie in the
    computed corner range widened by err;
  * invariants in ms and in the work cap;
  * degenerate cases that reduce to 2-D textbook tests;
  * planted mistakes (oracle mutations) that each of these catches.

The reference call sites are src/tools/tight_inclusion_full_query_benchmark
.cpp:375-397 with the parameters of :24-40 (ms 0, tolerance 1e-6, t_max 1,
max_itr 1e6)."""
import contextlib
import ctypes
from fractions import Fraction

import numpy as np
import pytest

import oracle

sp = pytest.importorskip("sympy")
import mpmath  # noqa: E402  (ships with sympy)

TOL = 1e-6


@contextlib.contextmanager
def mutation(k):
    m = ctypes.c_int.in_dll(oracle.lib(), "orc_tuv_mutation")
    m.value = k
    try:
        yield
    finally:
        m.value = 0


def _q(rows0, rows1):
    return np.array([list(rows0) + list(rows1)], dtype=np.float64)


# ---- numeric distances (float64; used only at the tolerance scale) --------
def _closest_on_triangle(p, a, b, c):
    # Ericson, Real-Time Collision Detection, 5.1.5
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + d1 / (d1 - d3) * ab
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return c
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + d2 / (d2 - d6) * ac
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (d4 - d3) / ((d4 - d3) + (d5 - d6)) * (c - b)
    denom = 1.0 / (va + vb + vc)
    return a + ab * (vb * denom) + ac * (vc * denom)


def _segment_distance(p1, q1, p2, q2):
    # Ericson 5.1.9, with the degenerate branches
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, e, f = d1 @ d1, d2 @ d2, d2 @ r
    if a <= 1e-300 and e <= 1e-300:
        return np.linalg.norm(r)
    if a <= 1e-300:
        s, t = 0.0, np.clip(f / e, 0, 1)
    else:
        c = d1 @ r
        if e <= 1e-300:
            t, s = 0.0, np.clip(-c / a, 0, 1)
        else:
            b = d1 @ d2
            den = a * e - b * b
            s = np.clip((b * f - c * e) / den, 0, 1) if den > 0 else 0.0
            t = (b * s + f) / e
            if t < 0:
                t, s = 0.0, np.clip(-c / a, 0, 1)
            elif t > 1:
                t, s = 1.0, np.clip((b - c) / a, 0, 1)
    return np.linalg.norm((p1 + d1 * s) - (p2 + d2 * t))


def _distance_at(is_ee, x, t):
    pts = x[:4] + t * (x[4:] - x[:4])
    if is_ee:
        return _segment_distance(pts[0], pts[1], pts[2], pts[3])
    return np.linalg.norm(pts[0] - _closest_on_triangle(pts[0], pts[1], pts[2], pts[3]))


# ---- exact first contact (sympy / mpmath) ---------------------------------
_T = sp.Symbol("t")


def _exact_first_contact(is_ee, x):
    """(t*, ok): t* the first contact time in [0,1] as an exact algebraic
    number (None if the primitives never meet); ok False when the case is
    degenerate or a contact lies within 1e-30 of a boundary (left to the
    degenerate-case tests)."""
    r0 = [[sp.Rational(Fraction(float(c))) for c in row] for row in x[:4]]
    r1 = [[sp.Rational(Fraction(float(c))) for c in row] for row in x[4:]]
    pts = [[r0[i][k] + _T * (r1[i][k] - r0[i][k]) for k in range(3)] for i in range(4)]
    if is_ee:
        a0, a1, b0, b1 = pts
        rows = [[a1[k] - a0[k] for k in range(3)], [b0[k] - a0[k] for k in range(3)],
                [b1[k] - a0[k] for k in range(3)]]
    else:
        p, f0, f1, f2 = pts
        rows = [[f1[k] - f0[k] for k in range(3)], [f2[k] - f0[k] for k in range(3)],
                [p[k] - f0[k] for k in range(3)]]
    poly = sp.Poly(sp.expand(sp.Matrix(rows).det()), _T)
    if poly.is_zero:
        return None, False
    mpmath.mp.dps = 60
    for root in poly.real_roots():
        if root < 0 or root > 1:
            continue
        tv = mpmath.mpf(str(sp.N(root, 70)))
        P = [[mpmath.mpf(str(r0[i][k])) + tv * (mpmath.mpf(str(r1[i][k])) - mpmath.mpf(str(r0[i][k])))
              for k in range(3)] for i in range(4)]
        if is_ee:
            d1 = [P[1][k] - P[0][k] for k in range(3)]
            d2 = [P[2][k] - P[3][k] for k in range(3)]
            rhs = [P[2][k] - P[0][k] for k in range(3)]
        else:
            d1 = [P[2][k] - P[1][k] for k in range(3)]
            d2 = [P[3][k] - P[1][k] for k in range(3)]
            rhs = [P[0][k] - P[1][k] for k in range(3)]
        dot = lambda a, b: sum(a[k] * b[k] for k in range(3))  # noqa: E731
        g11, g12, g22 = dot(d1, d1), dot(d1, d2), dot(d2, d2)
        det = g11 * g22 - g12 * g12
        if abs(det) < mpmath.mpf(10) ** -40:
            return None, False
        s = (g22 * dot(d1, rhs) - g12 * dot(d2, rhs)) / det
        w = (g11 * dot(d2, rhs) - g12 * dot(d1, rhs)) / det
        margins = [s, w, 1 - s, 1 - w] if is_ee else [s, w, 1 - s - w]
        if min(abs(m) for m in margins) < mpmath.mpf(10) ** -30:
            return None, False
        if min(margins) > 0:
            return root, True
    return None, True


def _random_cases(is_ee, n, seed):
    """Seeded cases aimed so that about half of them meet: the moving side
    passes the other near a parameter point drawn partly outside the face or
    segment, at a random time."""
    rng = np.random.default_rng(seed)
    grid = lambda *s: rng.integers(-16, 17, s) / 16.0  # noqa: E731  (exact in binary)
    out = []
    for _ in range(n):
        tc = rng.uniform(0.05, 0.95)
        a, b = rng.uniform(-0.4, 1.0, 2)
        if is_ee:
            e0 = grid(2, 3)
            e1 = e0 + grid(2, 3) * 0.25
            m = e0 + tc * (e1 - e0)
            target = m[0] + b * (m[1] - m[0])
            d = rng.normal(0, 1, 3)
            n_ = np.cross(m[1] - m[0], d)
            n_ /= np.linalg.norm(n_) + 1e-300
            h = rng.uniform(0.2, 1.0)
            c0 = target - a * d + tc * h * n_
            c1 = target - a * d - (1 - tc) * h * n_
            x0 = np.array([c0, c0 + d, e0[0], e0[1]])
            x1 = np.array([c1, c1 + d, e1[0], e1[1]])
        else:
            f0 = grid(3, 3)
            f1 = f0 + grid(3, 3) * 0.25
            m = f0 + tc * (f1 - f0)
            target = m[0] + a * (m[1] - m[0]) + b * (1 - a) * (m[2] - m[0])
            n_ = np.cross(m[1] - m[0], m[2] - m[0])
            n_ /= np.linalg.norm(n_) + 1e-300
            h = rng.uniform(0.2, 1.0)
            jit = rng.normal(0, 0.05, 3)
            x0 = np.array([target + tc * h * n_ + jit, *f0])
            x1 = np.array([target - (1 - tc) * h * n_ + jit, *f1])
        out.append(np.concatenate([x0, x1]))
    return np.array(out)


# ---- closed forms ------------------------------------------------------------
@pytest.mark.parametrize("z0,z1,px,py", [(1.0, -1.0, 0.25, 0.25), (0.3, -0.7, 0.1, 0.6), (0.9, -0.05, 0.4, 0.1),
                                         (1e-3, -1.0, 0.2, 0.2), (2.0, -0.1, 0.33, 0.33)])
def test_closed_form_vf_axis_crossing(z0, z1, px, py):
    tri = [(0, 0, 0), (1, 0, 0), (0, 1, 0)]
    x = _q([(px, py, z0)] + tri, [(px, py, z1)] + tri)
    r = oracle.tuv_cert(False, x)[0]
    exact = Fraction(z0) / (Fraction(z0) - Fraction(z1))
    assert r["hit"] == 1
    assert Fraction(float(r["toi"])) <= exact
    assert float(exact) - r["toi"] <= 2 * TOL / (z0 - z1)


@pytest.mark.parametrize("z0,z1,s", [(1.0, -1.0, 0.5), (0.25, -0.5, 0.125), (0.6, -0.01, 0.7)])
def test_closed_form_ee_axis_crossing(z0, z1, s):
    # edge a along y at x = s, moving down in z; edge b static along x at y = z = 0
    x = _q([(s, -1, z0), (s, 1, z0), (-1, 0, 0), (1, 0, 0)], [(s, -1, z1), (s, 1, z1), (-1, 0, 0), (1, 0, 0)])
    r = oracle.tuv_cert(True, x)[0]
    exact = Fraction(z0) / (Fraction(z0) - Fraction(z1))
    assert r["hit"] == 1
    assert Fraction(float(r["toi"])) <= exact
    assert float(exact) - r["toi"] <= 2 * TOL / (z0 - z1)


def _structural_misses():
    tri = [(0, 0, 0), (1, 0, 0), (0, 1, 0)]
    vf = [
        _q([(0.75, 0.75, 1)] + tri, [(0.75, 0.75, -1)] + tri),  # through the plane in the u + v > 1 half
        _q([(0.2, 0.2, 1)] + tri, [(0.2, 0.2, 1e-3)] + tri),    # stops 1e-3 above the face
        _q([(-0.1, 0.5, 1)] + tri, [(-0.1, 0.5, -1)] + tri),    # beside the face, 0.1 away
    ]
    ee = [
        _q([(0, 0, 1e-3), (1, 0, 1e-3), (0, -1, 0), (0, 1, 0)], [(0, 0, 1e-3), (1, 0, 1e-3), (0, -1, 0), (0, 1, 0)]),
        _q([(0.5, -1, 1), (0.5, 1, 1), (2, 0, 0), (3, 0, 0)], [(0.5, -1, -1), (0.5, 1, -1), (2, 0, 0), (3, 0, 0)]),
    ]
    return vf, ee


def test_structural_misses_report_no_contact():
    vf, ee = _structural_misses()
    for x in vf:
        assert oracle.tuv_cert(False, x)[0]["hit"] == 0
    for x in ee:
        assert oracle.tuv_cert(True, x)[0]["hit"] == 0


def test_coplanar_and_parallel_degenerate_cases():
    # in-plane motion: the 2-D answer is the time the point enters the triangle
    tri = [(0, 0, 0), (1, 0, 0), (0, 1, 0)]
    r = oracle.tuv_cert(False, _q([(-0.5, 0.25, 0)] + tri, [(0.5, 0.25, 0)] + tri))[0]
    assert r["hit"] == 1 and 0.5 - 2 * TOL <= r["toi"] <= 0.5
    r = oracle.tuv_cert(False, _q([(2, 2, 0)] + tri, [(1.5, 1.5, 0)] + tri))[0]
    assert r["hit"] == 0  # stays outside the triangle in its plane
    # parallel edges in one plane sliding into overlap at t = 0.5
    a0, a1 = (0, 0, 0), (1, 0, 0)
    r = oracle.tuv_cert(True, _q([a0, a1, (0.5, 1, 0), (1.5, 1, 0)], [a0, a1, (0.5, -1, 0), (1.5, -1, 0)]))[0]
    assert r["hit"] == 1 and 0.5 - 2 * TOL <= r["toi"] <= 0.5
    # parallel edges that never overlap along their line
    r = oracle.tuv_cert(True, _q([a0, a1, (2, 1, 0), (3, 1, 0)], [a0, a1, (2, -1, 0), (3, -1, 0)]))[0]
    assert r["hit"] == 0


# ---- exact brute force ---------------------------------------------------------
@pytest.mark.parametrize("is_ee", [False, True])
def test_exact_rational_brute_force(is_ee):
    cases = _random_cases(is_ee, 120, seed=11 + int(is_ee))
    res = oracle.tuv_cert(is_ee, cases)
    checked = contacts = 0
    for x, r in zip(cases, res):
        tstar, ok = _exact_first_contact(is_ee, x)
        if not ok:
            continue
        checked += 1
        if tstar is not None:
            contacts += 1
            assert r["hit"] == 1, "missed a contact"
            assert sp.Rational(Fraction(float(r["toi"]))) <= tstar, "reported a contact later than the first"
        if r["hit"]:
            # within the tolerance of touching at the reported time
            assert _distance_at(is_ee, x, r["toi"]) <= 4 * TOL
    assert checked >= 100 and 20 <= contacts <= checked - 20


# ---- the error bound -------------------------------------------------------------
def _exact_F(is_ee, x, t, u, v):
    X = [[Fraction(float(x[i][k])) + t * (Fraction(float(x[4 + i][k])) - Fraction(float(x[i][k]))) for k in range(3)]
         for i in range(4)]
    if is_ee:
        return [X[0][k] + u * (X[1][k] - X[0][k]) - (X[2][k] + v * (X[3][k] - X[2][k])) for k in range(3)]
    return [X[0][k] - (X[1][k] + u * (X[2][k] - X[1][k]) + v * (X[3][k] - X[1][k])) for k in range(3)]


def _enclosure_violations(is_ee, cases, seed):
    rng = np.random.default_rng(seed)
    L = oracle.lib()
    L.orc_tuv_box.argtypes = [ctypes.c_int, ctypes.c_void_p] + [ctypes.c_double] * 6 + [ctypes.c_void_p] * 3
    bad = 0
    for x in cases:
        x = np.ascontiguousarray(x)
        e = rng.integers(0, 20, 3)
        lo = [Fraction(int(rng.integers(0, 1 << int(k))), 1 << int(k)) for k in e]
        w = [Fraction(1, 1 << int(k)) for k in e]
        mn, mx, err = np.zeros(3), np.zeros(3), np.zeros(3)
        b = [float(lo[0]), float(lo[0] + w[0]), float(lo[1]), float(lo[1] + w[1]), float(lo[2]), float(lo[2] + w[2])]
        L.orc_tuv_box(int(is_ee), x.ctypes.data, *b, mn.ctypes.data, mx.ctypes.data, err.ctypes.data)
        pts = [(lo[0] + i * w[0], lo[1] + j * w[1], lo[2] + k * w[2]) for i in (0, 1) for j in (0, 1) for k in (0, 1)]
        pts += [(lo[0] + Fraction(int(rng.integers(1, 64)), 64) * w[0], lo[1] + Fraction(int(rng.integers(1, 64)), 64) * w[1],
                 lo[2] + Fraction(int(rng.integers(1, 64)), 64) * w[2]) for _ in range(4)]
        for t, u, v in pts:
            F = _exact_F(is_ee, x, t, u, v)
            for k in range(3):
                if not (Fraction(mn[k]) - Fraction(err[k]) <= F[k] <= Fraction(mx[k]) + Fraction(err[k])):
                    bad += 1
    return bad


@pytest.mark.parametrize("is_ee", [False, True])
def test_error_bound_encloses_exact_values(is_ee):
    from p2cccd import workloads

    qs = workloads.make_query_set(150, 150, seed=99)
    cases = qs.ee if is_ee else qs.vf
    assert _enclosure_violations(is_ee, cases, seed=3) == 0


# ---- invariants --------------------------------------------------------------
def test_invariants_in_ms_and_work_cap():
    from p2cccd import workloads

    qs = workloads.make_query_set(800, 800, seed=123)
    for is_ee, x in ((False, qs.vf), (True, qs.ee)):
        base = oracle.tuv_cert(is_ee, x)
        wide = oracle.tuv_cert(is_ee, x, ms=1e-3)
        assert np.all(wide["hit"] >= base["hit"])
        both = base["hit"] == 1
        assert np.all(wide["toi"][both] <= base["toi"][both])
        capped = oracle.tuv_cert(is_ee, x, max_itr=16, heap_cap=1 << 12)
        assert np.all(capped["hit"] >= base["hit"])
        assert np.all(capped["toi"][both] <= base["toi"][both])
        assert capped["capped"].sum() > 0


# ---- planted mistakes ---------------------------------------------------------
def test_planted_mistakes_are_caught():
    from p2cccd import workloads

    qs = workloads.make_query_set(60, 60, seed=99)
    with mutation(1):  # error bound dropped
        assert _enclosure_violations(False, qs.vf, seed=3) + _enclosure_violations(True, qs.ee, seed=3) > 0
    vf, _ = _structural_misses()
    with mutation(2):  # VF parameter domain taken as the whole unit square
        assert oracle.tuv_cert(False, vf[0])[0]["hit"] == 1
    with mutation(3):  # second EE edge built from the wrong vertex
        x = _q([(0.5, -1, 1), (0.5, 1, 1), (-1, 0, 0), (1, 0, 0)], [(0.5, -1, -1), (0.5, 1, -1), (-1, 0, 0), (1, 0, 0)])
        r = oracle.tuv_cert(True, x)[0]
        assert not (r["hit"] == 1 and 0.5 - 2 * TOL <= r["toi"] <= 0.5)
