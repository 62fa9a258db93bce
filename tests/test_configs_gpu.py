"""The benchmark configurations at full size on the GPU, against the oracle.

- Batched dense self-contact (256x256 folded cloth on a level-5 sphere,
  11.86 M candidate pairs): the graph-replayed batched step's hit and TOI for
  every pair equal the oracle's canonical certificate of that pair (the
  candidate set itself equals the oracle's sweep in test_broadphase_gpu.py).
- Learned group scheduling on the two level-6 icospheres: spatial cells of
  64-4096 pairs, the trained STPF priority on the MFMA kernel (against torch
  FP32), and each cell's first hit in its scheduled sequence and its decision
  equal to what oracle outcomes give, in scheduled and in plain order.
- Serving 32 resident scenes of about 1k-500k triangles under an HBM budget
  that forces evictions: every request's contacts are the same alone and in
  the mix, and sampled requests' candidate counts, contact pairs and TOIs
  equal oracle candidates plus oracle certificates.

The reference's all-exact rows are the model
(src/tools/scene_object_envelope_native_ti_walltime.cpp:1239-1281): every
candidate certified, zero false negatives."""
import dataclasses

import numpy as np
import pytest
import torch

import oracle
from p2cccd import scenes, schedule, serve, stpf, train_stpf
from p2cccd.pipeline import CcdPipeline
from p2cccd.tuv import TuvCertifier

pytestmark = pytest.mark.gpu

THREADS = 16   # host threads for the oracle (the CPU share of one GPU)
CHUNK = 1 << 21


def _positions(scene, kind, pairs, x0=None, x1=None):
    t = scene.topo
    x0 = scene.x0 if x0 is None else x0
    x1 = scene.x1 if x1 is None else x1
    if kind == "vf":
        vid = np.concatenate([pairs[:, :1], t.faces[pairs[:, 1]]], axis=1)
    else:
        vid = np.concatenate([t.edges[pairs[:, 0]], t.edges[pairs[:, 1]]], axis=1)
    return np.concatenate([x0[vid], x1[vid]], axis=1)


def _oracle_hits(scene, kind, pairs, x0=None, x1=None):
    hit = np.zeros(len(pairs), dtype=bool)
    toi = np.zeros(len(pairs))
    for a in range(0, len(pairs), CHUNK):
        o = oracle.tuv_cert(kind == "ee", _positions(scene, kind, pairs[a:a + CHUNK], x0, x1), threads=THREADS)
        hit[a:a + CHUNK] = o["hit"] != 0
        toi[a:a + CHUNK] = o["toi"]
    return hit, toi


def test_cloth_step_full_size_every_pair_matches_oracle(gpu):
    s = scenes.cloth_on_sphere()
    p = CcdPipeline(s, device=gpu)
    p.step_batched()
    p.step_batched()  # a graph replay of the captured step
    r = p.results()
    n = len(r["vf"]) + len(r["ee"])
    print(f"cloth step: {len(r['vf'])} VF + {len(r['ee'])} EE pairs", flush=True)
    assert n > 10_000_000 and r["replayed"] == 0
    contacts = 0
    for kind in ("vf", "ee"):
        res = r[kind + "_res"]
        hit, toi = _oracle_hits(s, kind, r[kind])
        assert np.array_equal(res["hit"] != 0, hit), kind            # identical decisions, no false negative
        assert np.array_equal(res["toi"], toi), kind                  # bit-identical TOI
        contacts += int(hit.sum())
    assert contacts > 0


def _first_hits(order, starts, sizes, hits):
    """First position of a hit in each group's sequence, -1 if none."""
    first = np.full(len(starts), -1, dtype=np.int64)
    pos = np.flatnonzero(hits[order])
    grp = np.searchsorted(starts, pos, side="right") - 1
    g, at = np.unique(grp, return_index=True)  # positions ascend: the first of each group
    first[g] = pos[at] - starts[g]
    return first


def test_config2_scene_learned_cell_schedule_matches_oracle(gpu):
    s = scenes.two_icospheres()  # two level-6 icospheres, 40,962 vertices each
    cert = TuvCertifier(device=gpu)
    packed = stpf.load_stpf(stpf.DEFAULT_PRIORITY_WEIGHTS, gpu)
    model = stpf.StpfPriority().to(gpu)
    with torch.no_grad():  # the same weights in torch FP32, the reference for the kernel
        w = packed
        at = 0
        for prm in (model.l1.weight, model.l1.bias, model.l2.weight, model.l2.bias, model.head.weight,
                    model.head.bias):
            prm.copy_(w[at:at + prm.numel()].view_as(prm))
            at += prm.numel()
    parts = train_stpf.scene_cells(cert, s)
    assert {d["kind"] for d in parts} == {"vf", "ee"}
    for d in parts:
        kind, cell, G = d["kind"], d["cell"], d["G"]
        sizes = torch.bincount(cell, minlength=G).cpu().numpy()
        assert sizes.max() <= 4096 and (sizes[:-1] >= 64).all() and sizes.sum() == d["pairs"].shape[0]
        prio = stpf.stpf_priority(packed, d["gfeat"])
        with torch.no_grad():
            assert torch.allclose(prio, model(d["gfeat"]), rtol=1e-4, atol=2e-6)
        pairs = d["pairs"].cpu().numpy()
        hits, _ = _oracle_hits(s, kind, pairs)
        truth = np.zeros(G, dtype=bool)
        np.logical_or.at(truth, cell.cpu().numpy(), hits)
        print(f"{kind}: {len(pairs)} pairs in {G} cells, {int(truth.sum())} with a contact", flush=True)
        assert truth.any()
        prox = d["feats"][:, 25]
        calls = {}
        for name, scores, top_k in (("scheduled", prox, 4096), ("plain", torch.zeros_like(prox), 0)):
            g = schedule.order_groups(cell, scores, top_k)
            res = schedule.any_hit(cert, kind, d["x0"], d["x1"], d["topo"], d["pairs"], g)
            assert np.array_equal(g.keys.cpu().numpy(), np.arange(G))
            first = _first_hits(g.order.cpu().numpy(), g.starts.cpu().numpy(), res.sizes, hits)
            assert np.array_equal(res.first_hit, first), (kind, name)
            assert np.array_equal(res.first_hit >= 0, truth), (kind, name)  # every cell decided exactly
            calls[name] = np.where(first >= 0, first + 1, res.sizes)
        # cells ranked by the learned priority reach their detections with fewer certificates on average
        sched = train_stpf._order_metrics(calls["scheduled"], truth, np.argsort(-prio.cpu().numpy(), kind="stable"))
        plain = train_stpf._order_metrics(calls["plain"], truth, np.arange(G))
        print(kind, "mean certificates at a detection: scheduled", sched["mean_detection"], "plain",
              plain["mean_detection"], flush=True)
        assert sched["mean_detection"] < plain["mean_detection"], kind


def test_serving_32_scenes_under_evictions_matches_alone_and_oracle(gpu):
    table, reqs = serve.poisson_workload(32, 48, 200.0, seed=9)
    tris = sorted(s.info["triangles"] for s in table.values())
    assert tris[0] <= 2000 and tris[-1] >= 400_000  # icosphere sizes come in steps of 4x
    est = {sid: serve.ResidentScene.estimate_bytes(s) for sid, s in table.items()}
    budget = int(sum(est.values()) * 0.35)
    assert budget > 2 * max(est.values())
    srv = serve.Server(table, budget)
    mixed = {c.rid: c for c in srv.run(reqs, realtime=False)}
    print(f"served {len(mixed)} requests in {srv.batches} batches, {srv.res.evictions} evictions", flush=True)
    assert len(mixed) == len(reqs) and srv.res.evictions > 0 and srv.batches >= 2
    contacts = 0
    for r in reqs:
        m = mixed[r.rid]
        alone = serve.Server({r.scene_id: table[r.scene_id]}, hbm_budget=4 * est[r.scene_id]).run(
            [serve.Request(r.rid, r.scene_id, r.x0, r.x1)], realtime=False)[0]
        assert (m.n_vf, m.n_ee) == (alone.n_vf, alone.n_ee), r.rid
        for a, b in zip((m.vf_hits, m.vf_toi, m.ee_hits, m.ee_toi),
                        (alone.vf_hits, alone.vf_toi, alone.ee_hits, alone.ee_toi)):
            assert np.array_equal(a, b), r.rid
        contacts += len(m.vf_hits) + len(m.ee_hits)
    assert contacts > 0
    # oracle: candidates by the sort-and-sweep, every candidate certified,
    # on the smallest, a middle and a large served scene
    by_size = sorted({r.scene_id for r in reqs}, key=lambda sid: table[sid].info["triangles"])
    for sid in (by_size[0], by_size[len(by_size) // 2], by_size[-4]):
        r = next(q for q in reqs if q.scene_id == sid)
        m = mixed[r.rid]
        sc = dataclasses.replace(table[sid], x0=r.x0, x1=r.x1)
        ref = oracle.candidates(sc)
        assert (m.n_vf, m.n_ee) == (len(ref["vf"]), len(ref["ee"])), sid
        for kind, got_pairs, got_toi in (("vf", m.vf_hits, m.vf_toi), ("ee", m.ee_hits, m.ee_toi)):
            hit, toi = _oracle_hits(sc, kind, ref[kind])
            want = ref[kind][hit]
            o = np.lexsort((got_pairs[:, 1], got_pairs[:, 0]))
            assert np.array_equal(got_pairs[o], want), (sid, kind)   # ref pairs are in lexicographic order
            assert np.array_equal(got_toi[o], toi[hit]), (sid, kind)
        print(f"scene {sid} ({table[sid].info['triangles']} triangles): {m.n_vf + m.n_ee} pairs, "
              f"{len(m.vf_hits) + len(m.ee_hits)} contacts, equal to the oracle", flush=True)
