"""Learned group scheduling: train the STPF priority scorer on certificate
outcomes the GPU path itself produces, then compare scheduled with plain order.

Groups are spatial cells of 64-4096 candidate pairs (schedule.spatial_cells);
each gets 32 FP32 features (schedule.group_features); its label is whether an
exact certificate finds a contact among its pairs.  Training data: seeded
variants of the plain query set (workloads.make_query_set) and of the two-
icosphere scene (scenes.two_icospheres) with varied start gap and rotation.
The model is the trunk and priority head of STPFModel
(src/python/p2cccd/proposal/stpf_model.py:45-93) at hidden 128; inference runs
through the hand-written kernel (stpf.stpf_priority).

Comparison on the held-out default two-icosphere scene: cells certified in
plain order (Morton cell order, pairs in list order) against scheduled order
(cells by learned priority, pairs by the proximity score with the reference's
top-K-then-fallback sequence), every cell certified until its first hit.
Exact certificates decide every cell in both orders; the decisions must be
identical, and the report gives how many certificates each order spends
before its detections.

    python -m p2cccd.train_stpf --out PATH.npz --report PATH.json
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np
import torch

from . import scenes, schedule, stpf, workloads
from .broadphase import BroadPhase
from .tuv import TUV_RESULT_DTYPE, TuvCertifier


def _certify_rows(cert: TuvCertifier, rows: torch.Tensor, is_ee: bool) -> torch.Tensor:
    n = rows.shape[0]
    cert.setup_positions(rows.contiguous(), is_ee)
    cert.certify(n, "stream")
    res = cert.result_tensor(n).view(torch.float64).view(n, 4)
    return (res[:, 2].view(torch.int64) & 0xFFFFFFFF) != 0


def cells_from_rows(cert: TuvCertifier, rows: torch.Tensor, kind: str):
    """Cells, their features and labels, and the per-pair data needed to
    schedule them."""
    n = rows.shape[0]
    idx = torch.arange(n, device=rows.device)
    f = stpf.tiny_features(kind, rows.contiguous(), idx)
    cell = schedule.spatial_cells(rows)
    G = int(cell.max()) + 1
    gf = schedule.group_features(rows, f, cell, G)
    hit = _certify_rows(cert, rows, kind == "ee")
    label = torch.zeros(G, dtype=torch.float32, device=rows.device)
    label.scatter_reduce_(0, cell, hit.float(), reduce="amax")
    return dict(rows=rows, feats=f, cell=cell, G=G, gfeat=gf, label=label, hit=hit)


def scene_cells(cert: TuvCertifier, scene) -> list[dict]:
    dev = torch.device("cuda")
    bp = BroadPhase.for_scene(scene, dev)
    x0, x1 = torch.from_numpy(scene.x0).to(dev), torch.from_numpy(scene.x1).to(dev)
    vf, ee = bp.candidates(x0, x1, scene.thickness)
    out = []
    for kind, pairs, topo in (("vf", vf, scene.topo.faces), ("ee", ee, scene.topo.edges)):
        if pairs.shape[0] < 64:
            continue
        t = torch.from_numpy(np.ascontiguousarray(topo)).to(dev)
        rows = schedule.candidate_rows(x0, x1, t, pairs, kind)
        d = cells_from_rows(cert, rows, kind)
        d.update(kind=kind, pairs=pairs, topo=t, x0=x0, x1=x1)
        out.append(d)
    return out


def query_cells(cert: TuvCertifier, seed: int) -> list[dict]:
    qs = workloads.make_query_set(20_000, 20_000, seed=seed)
    out = []
    for kind, x in (("vf", qs.vf), ("ee", qs.ee)):
        rows = torch.from_numpy(np.ascontiguousarray(x)).cuda()
        out.append(cells_from_rows(cert, rows, kind))
    return out


def collect(n_variants: int = 50, base_seed: int = workloads.BASE_SEED + 1000):
    cert = TuvCertifier()
    X, y = [], []
    rng = np.random.default_rng(base_seed)
    for i in range(n_variants):
        seed = base_seed + i
        if i % 2 == 0:
            parts = query_cells(cert, seed)
        else:
            gap = float(rng.uniform(0.0, 0.06))
            parts = scene_cells(cert, scenes.two_icospheres(level=6, gap=gap, seed=seed))
        for d in parts:
            X.append(d["gfeat"])
            y.append(d["label"])
        print(f"[train_stpf] variant {i}: {sum(int(d['G']) for d in parts)} cells", flush=True)
    return torch.cat(X), torch.cat(y)


def train(X: torch.Tensor, y: torch.Tensor, epochs: int = 300, lr: float = 2e-3, seed: int = 0):
    torch.manual_seed(seed)
    mu = X.mean(0)
    sd = X.std(0)
    # a feature that is constant over the training cells (the cell-count
    # column of ones, for one) normalises to 0 and carries nothing; dividing
    # by a floored std instead would fold into a huge first-layer weight and
    # an opposite huge bias, whose FP32 cancellation at inference would cost
    # more than the features themselves
    const = sd <= 1e-6 * X.abs().mean(0).clamp(min=1.0)
    sd = torch.where(const, torch.ones_like(sd), sd)
    Xn = (X - mu) / sd
    model = stpf.StpfPriority().to(X.device)
    opt = torch.optim.Adam(model.parameters(), lr=lr)
    pos = float(y.mean().clamp(min=1e-3, max=1 - 1e-3))
    weight = torch.tensor((1 - pos) / pos, device=X.device)
    lossf = torch.nn.BCEWithLogitsLoss(pos_weight=weight)
    for _ in range(epochs):
        perm = torch.randperm(Xn.shape[0], device=X.device)
        for b in range(0, Xn.shape[0], 4096):
            j = perm[b:b + 4096]
            opt.zero_grad()
            loss = lossf(model.logits(Xn[j]), y[j])
            loss.backward()
            opt.step()
    # fold the input normalisation into the first layer (in FP64) so the kernel takes raw features;
    # constant features get zero weight, which is what their normalised value 0 contributed
    with torch.no_grad():
        w = model.l1.weight.double()
        w[:, const] = 0.0
        W = w / sd.double()
        b = model.l1.bias.double() - (w * (mu.double() / sd.double())).sum(1)
        model.l1.weight.copy_(W.float())
        model.l1.bias.copy_(b.float())
    return model, float(loss.detach())


def _order_metrics(calls: np.ndarray, positive: np.ndarray, order: np.ndarray) -> dict:
    """Certificates spent when cells are processed in `order`, each until its
    first hit: total, before the first detection, and at each detection."""
    c = np.cumsum(calls[order])
    pos = positive[order]
    det = c[pos]
    return dict(total=int(c[-1]), first_detection=int(det[0]) if det.size else None,
                mean_detection=float(det.mean()) if det.size else None,
                median_detection=float(np.median(det)) if det.size else None)


def evaluate(packed: torch.Tensor, scene=None) -> dict:
    cert = TuvCertifier()
    scene = scene or scenes.two_icospheres()
    report = {"scene": scene.name, "kinds": {}}
    for d in scene_cells(cert, scene):
        kind = d["kind"]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        prio = stpf.stpf_priority(packed, d["gfeat"]).cpu().numpy()
        prox = d["feats"][:, 25]
        g_s = schedule.order_groups(d["cell"], prox, 4096)
        r_s = schedule.any_hit(cert, kind, d["x0"], d["x1"], d["topo"], d["pairs"], g_s)
        torch.cuda.synchronize()
        t_sched = time.perf_counter() - t0
        g_p = schedule.order_groups(d["cell"], torch.zeros_like(prox), 0)
        r_p = schedule.any_hit(cert, kind, d["x0"], d["x1"], d["topo"], d["pairs"], g_p)
        keys = g_s.keys.cpu().numpy()
        assert np.array_equal(keys, g_p.keys.cpu().numpy())
        pos_s, pos_p = r_s.first_hit >= 0, r_p.first_hit >= 0
        if not np.array_equal(pos_s, pos_p):
            raise AssertionError("scheduled and plain orders decided a cell differently")
        truth = d["label"].cpu().numpy()[keys] > 0
        if not np.array_equal(pos_s, truth):
            raise AssertionError("a cell with a certified contact was not detected")
        calls_s = np.where(pos_s, r_s.first_hit + 1, r_s.sizes)
        calls_p = np.where(pos_p, r_p.first_hit + 1, r_p.sizes)
        plain = _order_metrics(calls_p, pos_p, np.arange(len(keys)))
        sched = _order_metrics(calls_s, pos_s, np.argsort(-prio[keys], kind="stable"))
        report["kinds"][kind] = dict(cells=int(len(keys)), positive_cells=int(pos_s.sum()),
                                     pairs=int(d["pairs"].shape[0]), plain=plain, scheduled=sched,
                                     scheduled_certified=r_s.certified, rounds=r_s.rounds,
                                     schedule_seconds=t_sched, decisions_identical=True)
    return report


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--report", required=True)
    ap.add_argument("--variants", type=int, default=50)
    ap.add_argument("--epochs", type=int, default=300)
    args = ap.parse_args()
    t0 = time.perf_counter()
    X, y = collect(args.variants)
    print(f"[train_stpf] {X.shape[0]} cells, {float(y.mean()):.3f} positive, {time.perf_counter() - t0:.1f} s",
          flush=True)
    model, loss = train(X, y, args.epochs)
    print(f"[train_stpf] trained, last loss {loss:.4f}", flush=True)
    stpf.save_stpf(args.out, model, {"cells": X.shape[0], "positive_fraction": float(y.mean())})
    rep = evaluate(model.packed().cuda())
    rep.update(training_cells=int(X.shape[0]), positive_fraction=float(y.mean()), variants=args.variants)
    with open(args.report, "w") as fh:
        json.dump(rep, fh, indent=1)
    print(json.dumps(rep), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
