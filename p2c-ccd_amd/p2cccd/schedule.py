"""Scheduled any-hit certification of candidate groups on the GPU.

The reference's scheduled rows (src/tools/scene_object_envelope_native_ti_walltime
.cpp:1283-1410, RunScheduledAnyHitRows) certify, per group, the top-K
candidates by score in descending order and stop at the first hit; a group
without a hit then replays its remaining candidates in list order until one
hits (the fallback, :1367-1398).  Per group that is one sequence -- top-K by
score, then the rest by index -- certified until its first hit.  Scores only
order the work: whether a group holds a hit is decided by exact certificates
and never by a score, so the detected groups equal the all-exact ones.

On the device the sequences are certified in rounds: round r certifies the
next c_r candidates of every unresolved group together (c_r = 1, 2, 4, ...),
one batched certificate launch per round, then each group keeps its first hit
position.  The result per group is the reference's (first hit position in the
sequence, hence its exact-call count); the rounds only add the overshoot of
the last chunk as extra, discarded work, which is reported separately.

Ties in score are broken by candidate index (the reference's std::sort leaves
their order unspecified).

Groups are either object pairs (GroupKey, :756-771, from the TinySTPF scorer)
or spatial cells: runs of candidates in Morton order of their swept-box
centres, cut to 64-4096 pairs, with 32 FP32 features per cell for the STPF
priority scorer.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .tuv import TUV_RESULT_DTYPE, TuvCertifier


@dataclass
class GroupOrder:
    order: torch.Tensor    # candidate indices, grouped, each group in its certification order
    starts: torch.Tensor   # int64[G] first position of each group in `order`
    sizes: torch.Tensor    # int64[G]
    keys: torch.Tensor     # int64[G] group key
    top: torch.Tensor      # int64[G] candidates in each group's scored (top-K) part


def order_groups(keys: torch.Tensor, scores: torch.Tensor, top_k: int) -> GroupOrder:
    """Per group (ascending key): the top_k candidates by descending score
    (ties: ascending index), then the rest by ascending index."""
    n = keys.numel()
    dev = keys.device
    idx = torch.arange(n, device=dev)
    o = torch.argsort(-scores, stable=True)
    o = o[torch.argsort(keys[o], stable=True)]
    ks = keys[o]
    new = torch.ones(n, dtype=torch.bool, device=dev)
    new[1:] = ks[1:] != ks[:-1]
    starts = torch.nonzero(new).squeeze(1)
    sizes = torch.diff(torch.cat([starts, torch.tensor([n], device=dev)]))
    gid = torch.cumsum(new.long(), 0) - 1
    rank = idx - starts[gid]
    phase = (rank >= top_k).long()
    second = torch.where(phase.bool(), o, rank)
    # sequence key inside the group: (phase, rank or index)
    s = torch.argsort(second, stable=True)
    s = s[torch.argsort(phase[s], stable=True)]
    s = s[torch.argsort(gid[s], stable=True)]
    return GroupOrder(order=o[s], starts=starts, sizes=sizes, keys=ks[starts], top=torch.clamp(sizes, max=top_k))


@dataclass
class AnyHitResult:
    first_hit: np.ndarray      # int64[G]: position of the first hit in the group's sequence, -1 if none
    sizes: np.ndarray          # int64[G]
    exact_calls: int           # the reference's count: first hit + 1, or the group size
    certified: int             # certificates actually run (chunk overshoot included)
    rounds: int
    toi: np.ndarray            # float64[G]: TOI of the first hit (inf if none)


def any_hit(cert: TuvCertifier, kind: str, x0: torch.Tensor, x1: torch.Tensor, topo: torch.Tensor,
            pairs: torch.Tensor, g: GroupOrder, chunk0: int = 1, growth: int = 2, mode: str = "stream") -> AnyHitResult:
    """Certifies each group's sequence until its first hit (see the module doc)."""
    dev = pairs.device
    G = g.starts.numel()
    pos = torch.zeros(G, dtype=torch.int64, device=dev)
    first = torch.full((G,), -1, dtype=torch.int64, device=dev)
    toi = torch.full((G,), float("inf"), dtype=torch.float64, device=dev)
    open_ = torch.ones(G, dtype=torch.bool, device=dev)
    chunk, certified, rounds = chunk0, 0, 0
    is_ee = kind == "ee"
    while True:
        length = torch.where(open_, torch.clamp(g.sizes - pos, min=0, max=chunk), torch.zeros_like(pos))
        total = int(length.sum())
        if total == 0:
            break
        grp = torch.repeat_interleave(torch.arange(G, device=dev), length)
        offs = torch.cumsum(length, 0) - length
        within = torch.arange(total, device=dev) - offs[grp]
        seqpos = pos[grp] + within
        cand = g.order[g.starts[grp] + seqpos]
        batch = pairs[cand].contiguous()
        cert.reserve(total)
        cert.setup_pairs(x0, x1, topo, batch, is_ee, total)
        cert.certify(total, mode)
        res = cert.result_tensor(total).view(torch.float64).view(total, 4)
        hit = res[:, 2].view(torch.int64) & 0xFFFFFFFF
        hitpos = torch.where(hit != 0, seqpos, torch.full_like(seqpos, 1 << 62))
        best = torch.full((G,), 1 << 62, dtype=torch.int64, device=dev)
        best.scatter_reduce_(0, grp, hitpos, reduce="amin")
        found = open_ & (best < (1 << 62))
        # TOI of the first hit of each newly resolved group
        sel = (hit != 0) & (seqpos == best[grp])
        toi_items = torch.where(sel, res[:, 0], torch.full_like(res[:, 0], float("inf")))
        t = torch.full((G,), float("inf"), dtype=torch.float64, device=dev)
        t.scatter_reduce_(0, grp, toi_items, reduce="amin")
        first = torch.where(found, best, first)
        toi = torch.where(found, t, toi)
        open_ = open_ & ~found
        pos = pos + length
        open_ = open_ & (pos < g.sizes)
        certified += total
        rounds += 1
        chunk *= growth
    first_np = first.cpu().numpy()
    sizes = g.sizes.cpu().numpy()
    exact = int(np.where(first_np >= 0, first_np + 1, sizes).sum())
    return AnyHitResult(first_hit=first_np, sizes=sizes, exact_calls=exact, certified=certified, rounds=rounds,
                        toi=toi.cpu().numpy())


# ---- spatial cells ----------------------------------------------------------
def _spread10(v: torch.Tensor) -> torch.Tensor:
    v = v & 0x3FF
    v = (v | (v << 16)) & 0x030000FF
    v = (v | (v << 8)) & 0x0300F00F
    v = (v | (v << 4)) & 0x030C30C3
    v = (v | (v << 2)) & 0x09249249
    return v


def candidate_rows(x0: torch.Tensor, x1: torch.Tensor, topo: torch.Tensor, pairs: torch.Tensor, kind: str):
    """(n, 8, 3) positions of each candidate's 4 vertices at t0 then t1."""
    if kind == "vf":
        vid = torch.cat([pairs[:, :1], topo[pairs[:, 1].long()]], dim=1).long()
    else:
        vid = torch.cat([topo[pairs[:, 0].long()], topo[pairs[:, 1].long()]], dim=1).long()
    return torch.cat([x0[vid], x1[vid]], dim=1)


def spatial_cells(rows: torch.Tensor, min_size: int = 64, max_size: int = 4096, level: int = 5) -> torch.Tensor:
    """Cell id per candidate: candidates sorted by the 30-bit Morton code of
    their swept-box centre, cut where the level-`level` Morton prefix changes
    (or at max_size), runs shorter than min_size merged into the next run."""
    mn = rows.amin(dim=1)
    mx = rows.amax(dim=1)
    c = 0.5 * (mn + mx)
    lo, hi = c.amin(dim=0), c.amax(dim=0)
    q = ((c - lo) / torch.clamp(hi - lo, min=1e-300) * 1023.0).clamp(0, 1023).long()
    code = (_spread10(q[:, 0]) << 2) | (_spread10(q[:, 1]) << 1) | _spread10(q[:, 2])
    order = torch.argsort(code, stable=True)
    prefix = (code[order] >> (30 - 3 * level)).cpu().numpy()
    n = prefix.size
    # cut points: prefix changes, then split long runs and merge short ones
    change = np.flatnonzero(np.diff(prefix)) + 1
    bounds = np.concatenate([[0], change, [n]])
    cuts = [0]
    for a, b in zip(bounds[:-1], bounds[1:]):
        start = cuts[-1]
        # close the pending run at a when it is long enough
        if a - start >= min_size and a != start:
            cuts.append(a)
            start = a
        while b - start > max_size:
            start += max_size
            cuts.append(start)
    if cuts[-1] != n:
        cuts.append(n)
    cuts = np.asarray(cuts)
    if len(cuts) > 2 and cuts[-1] - cuts[-2] < min_size:
        cuts = np.delete(cuts, -2)  # merge a short tail into the previous cell
    cell_sorted = np.repeat(np.arange(len(cuts) - 1), np.diff(cuts))
    cell = torch.empty(n, dtype=torch.int64, device=rows.device)
    cell[order] = torch.from_numpy(cell_sorted).to(rows.device)
    return cell


def group_features(rows: torch.Tensor, cand_feats: torch.Tensor, cell: torch.Tensor, n_cells: int,
                   tolerance: float = 1e-6) -> torch.Tensor:
    """32 FP32 features per cell: pair count, cell box extents, relative
    velocity, gap lower bounds and depth estimates, as sums / minima / maxima
    over its candidates' TinySTPF features (float64 [n, 32])."""
    dev = rows.device
    f = cand_feats
    G = n_cells

    def red(v, how):
        init = {"amin": float("inf"), "amax": float("-inf"), "sum": 0.0}[how]
        out = torch.full((G,), init, dtype=torch.float64, device=dev)
        return out.scatter_reduce_(0, cell, v.double(), reduce=how, include_self=True)

    cnt = red(torch.ones_like(f[:, 0]), "sum")
    mean = lambda v: red(v, "sum") / cnt  # noqa: E731
    mn = torch.stack([red(rows[:, :, k].amin(1), "amin") for k in range(3)], 1)
    mx = torch.stack([red(rows[:, :, k].amax(1), "amax") for k in range(3)], 1)
    ext = mx - mn
    gap_lb = red(f[:, 11], "amin")
    depth = torch.log2(torch.clamp(f[:, 5], min=tolerance) / tolerance)
    cols = [
        torch.log1p(cnt), ext[:, 0], ext[:, 1], ext[:, 2], torch.linalg.norm(ext, dim=1),
        mean(f[:, 7]), red(f[:, 7], "amax"), mean(f[:, 8]), red(f[:, 8], "amax"),
        gap_lb, mean(f[:, 11]), red(f[:, 9], "amin"), red(f[:, 10], "amin"),
        red(f[:, 25], "amax"), mean(f[:, 25]), mean(f[:, 26]), red(f[:, 26], "amax"),
        mean(depth), red(depth, "amax"), red(depth, "amin"),
        mean(f[:, 5]), red(f[:, 5], "amax"), mean(f[:, 17]), mean(f[:, 15]), red(f[:, 15], "amax"),
        mean(f[:, 0]), mean(f[:, 27]), mean(f[:, 28]), red(f[:, 28], "amax"),
        mean(f[:, 29]), torch.log1p(red(f[:, 8], "amax") / torch.clamp(gap_lb, min=1e-12)), torch.ones_like(cnt),
    ]
    return torch.stack(cols, 1).float()


def results_hits(cert: TuvCertifier, n: int) -> np.ndarray:
    r = cert.result_tensor(n).cpu().numpy().view(TUV_RESULT_DTYPE)
    return r["hit"]
