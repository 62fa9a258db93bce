"""Seeded synthetic workloads for the certificate path.

All generators use numpy PCG64 with a fixed base seed, so every run, test and
benchmark sees the same inputs.  Query layout: (n, 8, 3) float64, rows 0-3
are the four vertices at t=0 (VF: point, f0, f1, f2; EE: a0, a1, b0, b1) and
rows 4-7 the same vertices at t=1, the argument order of the reference's
Tight-Inclusion call sites (src/tools/tight_inclusion_full_query_benchmark.cpp:375-397).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

BASE_SEED = 20261203


@dataclass
class QuerySet:
    vf: np.ndarray          # (n_vf, 8, 3)
    ee: np.ndarray          # (n_ee, 8, 3)
    vf_kind: np.ndarray     # 0 random, 1 touching, 2 degenerate
    ee_kind: np.ndarray


def _unit(v: np.ndarray) -> np.ndarray:
    n = np.linalg.norm(v, axis=-1, keepdims=True)
    return v / np.where(n > 0, n, 1.0)


def _touch_vf(x: np.ndarray, rng: np.random.Generator, gap: float) -> None:
    """Shift the point's trajectory so that at a random t* it sits `gap` above
    a random interior point of the moving triangle."""
    n = x.shape[0]
    ts = rng.uniform(0.0, 1.0, n)
    r1, r2 = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
    flip = r1 + r2 > 1.0
    r1[flip], r2[flip] = 1.0 - r1[flip], 1.0 - r2[flip]
    at = (1.0 - ts)[:, None, None] * x[:, 0:4] + ts[:, None, None] * x[:, 4:8]
    a, b, c = at[:, 1], at[:, 2], at[:, 3]
    q = a + r1[:, None] * (b - a) + r2[:, None] * (c - a)
    nrm = _unit(np.cross(b - a, c - a))
    target = q + gap * nrm
    shift = target - at[:, 0]
    x[:, 0] += shift
    x[:, 4] += shift


def _touch_ee(x: np.ndarray, rng: np.random.Generator, gap: float) -> None:
    n = x.shape[0]
    ts = rng.uniform(0.0, 1.0, n)
    s, r = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
    at = (1.0 - ts)[:, None, None] * x[:, 0:4] + ts[:, None, None] * x[:, 4:8]
    pa = at[:, 0] + s[:, None] * (at[:, 1] - at[:, 0])
    pb = at[:, 2] + r[:, None] * (at[:, 3] - at[:, 2])
    nrm = _unit(np.cross(at[:, 1] - at[:, 0], at[:, 3] - at[:, 2]))
    shift = pa + gap * nrm - pb
    for k in (2, 3, 6, 7):
        x[:, k] += shift


def _degenerate_vf(x: np.ndarray, rng: np.random.Generator) -> None:
    """Project all 8 positions onto one random plane: point and triangle stay
    coplanar for every t."""
    n = x.shape[0]
    nrm = _unit(rng.normal(size=(n, 3)))
    off = rng.uniform(-0.5, 0.5, n)
    d = np.einsum("nkj,nj->nk", x, nrm) - off[:, None]
    x -= d[:, :, None] * nrm[:, None, :]


def _degenerate_ee(x: np.ndarray, rng: np.random.Generator) -> None:
    """Make edge b parallel to edge a at t=0 and t=1."""
    lam = rng.uniform(0.5, 1.5, (x.shape[0], 1))
    x[:, 3] = x[:, 2] + lam * (x[:, 1] - x[:, 0])
    x[:, 7] = x[:, 6] + lam * (x[:, 5] - x[:, 4])


def make_query_set(n_vf: int = 10_000, n_ee: int = 10_000, seed: int = BASE_SEED,
                   touch_frac: float = 0.10, degenerate_frac: float = 0.05,
                   gap: float = 1.0e-8, sigma: float = 0.1) -> QuerySet:
    """Plain CCD query set: starts U[-1,1]^3, displacement N(0, sigma) per
    axis; `touch_frac` constructed to pass within `gap` at a random t;
    `degenerate_frac` coplanar (VF) or parallel-edge (EE)."""
    rng = np.random.Generator(np.random.PCG64(seed))

    def one(n, touch, degen):
        x0 = rng.uniform(-1.0, 1.0, (n, 4, 3))
        x1 = x0 + rng.normal(0.0, sigma, (n, 4, 3))
        x = np.concatenate([x0, x1], axis=1)
        kind = np.zeros(n, np.int8)
        perm = rng.permutation(n)
        nt, nd = int(round(touch_frac * n)), int(round(degenerate_frac * n))
        ti, di = np.sort(perm[:nt]), np.sort(perm[nt:nt + nd])
        kind[ti], kind[di] = 1, 2
        xt, xd = x[ti], x[di]
        touch(xt, rng, gap)
        degen(xd, rng)
        x[ti], x[di] = xt, xd
        return x, kind

    vf, vk = one(n_vf, _touch_vf, _degenerate_vf)
    ee, ek = one(n_ee, _touch_ee, _degenerate_ee)
    return QuerySet(vf=vf, ee=ee, vf_kind=vk, ee_kind=ek)


def exact_queue(n: int, pool: int = 10_000, seed: int = BASE_SEED):
    """An exact work queue of n entries (p2cccd.exact_queue.QUEUE_ENTRY_DTYPE)
    over pools of point-triangle and edge-edge primitives taken from the plain
    query set: masks 1, 2 and 3, full and partial intervals, 0-4 primitives
    of each family per entry, the reference's default certificate config
    (eps_time 1e-4, eps_space 1e-6, depth 32).  Returns (entries, pts, ees)."""
    from . import _native
    from .exact_queue import QUEUE_ENTRY_DTYPE

    rng = np.random.Generator(np.random.PCG64(seed + 11))
    qs = make_query_set(pool, pool, seed=seed)
    pts = _native.prims_from_positions(qs.vf)
    ees = _native.prims_from_positions(qs.ee)
    e = np.zeros(n, QUEUE_ENTRY_DTYPE)
    e["work_item_id"] = np.arange(1, n + 1, dtype=np.uint64)
    e["parent_candidate_id"] = rng.integers(1, 1 << 40, n).astype(np.uint64)
    e["query_id"] = rng.integers(1, max(2, n // 4), n).astype(np.uint64)
    e["feature_family_mask"] = rng.choice([1, 2, 3], n, p=[0.4, 0.4, 0.2])
    t = np.sort(rng.uniform(0.0, 1.0, (n, 2)), axis=1)
    full = rng.uniform(size=n) < 0.5
    t[full] = (0.0, 1.0)
    e["interval_t0"], e["interval_t1"] = t[:, 0], t[:, 1]
    e["max_subdivision_depth"] = 32
    e["eps_time"] = 1.0e-4
    e["eps_space"] = 1.0e-6
    for fam in ("pt", "ee"):
        count = rng.integers(0, 5, n)
        e[f"{fam}_count"] = count
        e[f"{fam}_begin"] = rng.integers(0, pool - count + 1)
    return e, pts, ees
