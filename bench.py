"""Benchmark: certified VF/EE pairs per second for one CCD timestep on one
MI355X, with p50/p99 step latency.

Workload (the folded-cloth self-contact scene): a 256 x 256 cloth (65,536 V,
130,050 F) folded over itself with N(0, 1e-3) noise and per-vertex
velocities N(0, 0.02), falling onto a noisy level-5 sphere (20,480 F);
FP64 positions resident in HBM.  A step = LBVH broad phase + the (t,u,v)
inclusion certificate (tolerance 1e-6, ms 0, max_itr 1e6) on every VF and EE
candidate, fallback replay included, so every candidate ends with a final
decision.  The headline runs the batched form (device-side counts, stream
certificate kernel and wave-cooperative pass, HIP-graph replay); the plain
one-launch-per-stage form is timed in the same run.

Prints one JSON line.  `python bench.py --gpus N`: only N = 1 is supported.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT), str(ROOT / "p2c-ccd_amd")]

METRIC = "certified VF/EE pairs/sec + p50/p99 step latency, zero false negatives"
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector, FMA counted as 2
HBM_PEAK_GBS = 8000.0
# FP64 add/sub/mul per box evaluation (cert_tuv_math.h, tuv_classify_at):
# 3 upper corners, then 3 coordinates x 2 t-planes x (8 lerp + 2 edge + 4 + 2
# + 8) VF or (8 + 2 + 4 + 4 + 4) EE, plus 3 range widths.  Min/max and
# comparisons are not counted.
FLOP_BOX_VF = 3 + 3 * 2 * 24 + 3
FLOP_BOX_EE = 3 + 3 * 2 * 22 + 3


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def profiled_traffic(kernel: str):
    """HBM bytes per launch of `kernel` (read + written) from the committed
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE summary of this workload, scaled
    by a 1 GiB device copy; None when the summary has no line for it."""
    f = Path(__file__).resolve().parent / "profiles" / "hbm_traffic_cloth256_step.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        w = line.split()
        if len(w) == 4 and w[0] == "traffic_bytes_per_launch" and w[1] == kernel:
            return int(w[2].split("=")[1]) + int(w[3].split("=")[1])
    return None


def stream_roofline(pipe, steps: int):
    """The stream kernel's algorithmic FP64 rate over `steps` event-timed
    batched steps: boxes it evaluated (every pair's boxes minus what the wave
    pass added to the walks it took over) x FLOPs per box, over its average
    duration.  Returns (roofline dict, certificate-stage ms list, pairs handed
    to the wave pass)."""
    import numpy as np
    import torch

    stream_ms, stage_ms = [], []
    for _ in range(steps):
        ev = tuple(torch.cuda.Event(enable_timing=True) for _ in range(3))
        pipe.step_batched_timed(ev)
        stream_ms.append(ev[0].elapsed_time(ev[2]))
        stage_ms.append(ev[0].elapsed_time(ev[1]))
    rt = pipe.results()
    n_vf = pipe.last.n_vf
    boxes = np.concatenate([rt["vf_res"]["boxes"], rt["ee_res"]["boxes"]]).astype(np.int64)
    cert = pipe.cert
    hc = cert.heavy_count()
    heavy = cert.heavy_list[:hc].long()
    # the box count a handed-over walk had reached (WalkState.boxes, byte 32)
    saved = cert.states.view(-1, cert.state_bytes)[heavy][:, 32:36].contiguous().view(torch.int32)
    heavy = heavy.cpu().numpy()
    wave_part = np.maximum(boxes[heavy] - saved.cpu().numpy()[:, 0].astype(np.int64), 0)
    is_vf = np.arange(len(boxes)) < n_vf
    stream_vf = int(boxes[is_vf].sum() - wave_part[heavy < n_vf].sum())
    stream_ee = int(boxes[~is_vf].sum() - wave_part[heavy >= n_vf].sum())
    flops = stream_vf * FLOP_BOX_VF + stream_ee * FLOP_BOX_EE
    avg = float(np.mean(stream_ms))
    achieved = flops / (avg * 1e-3) / 1e12
    return ({"bound": "fp64-valu", "kernel": "tuv_stream_kernel", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
             "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None, "avg_ms": avg,
             "flops_per_launch": flops, "boxes_per_launch": stream_vf + stream_ee}, stage_ms, hc)


def broadphase_roofline(pipe, steps: int) -> dict:
    """The broad phase (build, count, emit) against HBM: algorithmic bytes of
    one step over its event-timed duration.  Bytes per step, with V vertices,
    F faces, E edges, B = V + F + E swept boxes and C candidate pairs:
      V * 48 (x0, x1 read) + B * 48 (boxes written) + B * 8 * 2 * 4 (u32 key
      and value read and written by each of 4 sort passes) + (2F - 1 + 2E - 1)
      * 64 (nodes written) + visited * 64 (one node read per visit of the
      count and emit walks, counted by a host replay of the walks) + C * 8
      (pairs written)."""
    import numpy as np
    import torch

    bp = pipe.bp
    ms = []
    for _ in range(steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        bp.build(pipe.x0, pipe.x1, pipe.thickness)
        bp.count()
        bp.emit(pipe.vf, pipe.ee)
        e1.record()
        torch.cuda.synchronize(pipe.device)
        ms.append(e0.elapsed_time(e1))
    visits = bp.traversal_visits()
    nv, nf, ne = bp.n_vertices, bp.n_faces, bp.n_edges
    n_vf, n_ee, _ = (int(v) for v in bp.totals.cpu())
    boxes = nv + nf + ne
    walked = sum(int(v) for v in visits.values())
    traffic = (nv * 48 + boxes * 48 + boxes * 8 * 2 * 4 + (max(2 * nf - 1, 0) + max(2 * ne - 1, 0)) * 64
               + walked * 64 + (n_vf + n_ee) * 8)
    avg = float(np.mean(ms))
    achieved = traffic / (avg * 1e-3) / 1e9
    return {"bound": "hbm", "stage": "broad phase (build + count + emit)", "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "avg_ms": avg, "visited_nodes": {k: int(v) for k, v in visits.items()}, "candidates": n_vf + n_ee}


def percentile(v, q):
    import numpy as np

    return float(np.percentile(np.asarray(v), q))


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, sample_pairs: int, seed: int = 7) -> dict:
    """The oracle (a restatement of the reference's CPU path: sort-and-sweep
    candidates, then the certificate per candidate) on one host thread, as
    the reference runs.  Sample: the full sweep over the scene, plus the
    certificate on a uniform random sample of its candidates, scaled to all."""
    import numpy as np

    import oracle

    t0 = time.perf_counter()
    cand = oracle.candidates(scene)
    t_sweep = time.perf_counter() - t0
    n_total = len(cand["vf"]) + len(cand["ee"])
    rng = np.random.default_rng(seed)
    t_cert = 0.0
    sampled = 0
    samples = []
    for kind in ("vf", "ee"):
        pairs = cand[kind]
        k = int(round(sample_pairs * len(pairs) / max(n_total, 1)))
        idx = rng.choice(len(pairs), size=min(k, len(pairs)), replace=False)
        t = scene.topo
        sel = pairs[idx]
        if kind == "vf":
            vid = np.concatenate([sel[:, :1], t.faces[sel[:, 1]]], axis=1)
        else:
            vid = np.concatenate([t.edges[sel[:, 0]], t.edges[sel[:, 1]]], axis=1)
        x = np.concatenate([scene.x0[vid], scene.x1[vid]], axis=1)
        scale = len(pairs) / max(len(idx), 1)
        samples.append((kind, x, scale))
        t1 = time.perf_counter()
        oracle.tuv_cert(kind == "ee", x, threads=1)
        t_cert += (time.perf_counter() - t1) * scale
        sampled += len(idx)
    step_s = t_sweep + t_cert
    # the same sample on every host core (OpenMP over the sampled candidates; the sweep stays one thread)
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    t_all = 0.0
    for kind, x, scale in samples:
        t1 = time.perf_counter()
        oracle.tuv_cert(kind == "ee", x, threads=cores)
        t_all += (time.perf_counter() - t1) * scale
    return {
        "value": n_total / step_s,
        "unit": "pairs/s",
        "cores": 1,
        "kind": "port",
        "sample": f"full sort-and-sweep of the scene ({n_total} candidates, {t_sweep:.1f} s) + certificate on "
                  f"{sampled} uniformly sampled candidates scaled to all ({t_cert:.1f} s est.), 1 thread",
        "all_cores": {"value": n_total / (t_sweep + t_all), "cores": cores, "nproc": os.cpu_count(),
                      "cpu_model": _cpu_model(),
                      "sample": f"same sample, certificate on {cores} threads ({t_all:.1f} s est.), sweep 1 thread"},
    }


def probe(pipe, steps: int) -> int:
    """Batched steps outside the graph (one dispatch per kernel, so counters
    attach to single launches), then a 1 GiB device-to-device copy whose
    algorithmic bytes are known (1 GiB read + 1 GiB written)."""
    import torch

    pipe.step_batched(use_graph=False)
    ms = []
    for _ in range(steps):
        ev = tuple(torch.cuda.Event(enable_timing=True) for _ in range(3))
        pipe.step_batched_timed(ev)
        ms.append(ev[0].elapsed_time(ev[2]))
    a = torch.empty(1 << 28, dtype=torch.float32, device=pipe.device)
    a.fill_(1.0)
    c = torch.empty_like(a)
    torch.cuda.synchronize(pipe.device)
    c.copy_(a)
    torch.cuda.synchronize(pipe.device)
    print(json.dumps({"probe": "stream kernel ms per step, then a 1 GiB copy", "stream_ms": ms,
                      "pairs": pipe.last.pairs, "copy_bytes": 2 * a.numel() * 4}), flush=True)
    return 0


FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 (dense)
STPF_FLOPS_ROW = 2 * (32 * 128 + 128 * 128 + 128)      # three affine layers, multiply-add counted as 2
TINY_FLOPS_ROW = 2 * (32 * 64 + 64 * 32 + 32) + 2 * 32  # TinySTPF MLP + z-score


def _gpu_ms(fn, repeats: int = 5) -> float:
    """Best-of-`repeats` device time of fn() on the current stream (HIP events)."""
    import torch

    best = float("inf")
    for _ in range(repeats):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    return best


def _cpu_s(fn) -> float:
    t = time.perf_counter()
    fn()
    return time.perf_counter() - t


def _row(name, workload, gpu, cpu, roofline=None, parity=None, extra=None):
    line = {"row": name, "workload": workload, "gpu": gpu, "roofline": roofline, "cpu_baseline": cpu,
            "parity": parity}
    line.update(extra or {})
    print(json.dumps(line), flush=True)


def rows_suite() -> int:
    """One JSON line per hot-path row: the GPU entry point on its workload,
    its roofline where the work has a closed-form count, and the reference's
    CPU path (compiled in place where it builds, else the oracle restatement)
    timed on this host, one thread, with a parity check of the two outputs."""
    import ctypes

    import numpy as np
    import torch

    import oracle
    from p2cccd import _native as nat
    from p2cccd import proxy, scenes, serve, stpf, workloads
    from p2cccd.pipeline import CcdPipeline
    from p2cccd.tuv import TuvCertifier

    dev = torch.device("cuda:0")
    qs = workloads.make_query_set(10_000, 10_000)

    # (t,u,v) certificate on the plain query set (10k VF + 10k EE)
    log("rows: (t,u,v) certificate, 10k VF + 10k EE queries")
    import tempfile

    from p2cccd.autotune import CertificateTuner, summarize

    for kind, x in (("vf", qs.vf), ("ee", qs.ee)):
        cert = TuvCertifier(device=dev)
        xd = torch.from_numpy(np.ascontiguousarray(x)).to(dev)
        n = cert.setup_positions(xd, kind == "ee")
        # the same autotuner as the headline, on these pairs (every candidate must reproduce the plain kernel)
        with tempfile.TemporaryDirectory() as td:
            _, trials = CertificateTuner(cert, n, cache=os.path.join(td, "tune.json")).tune(use_cache=False)
        tuned = {k: getattr(cert.cfg, k) for k in ("refill_at", "light_budget", "handover_at", "blocks_per_cu",
                                                   "split_at")}
        tuned.update(summarize(trials))
        res = cert.certify_positions(xd, kind == "ee")
        ms = {m: _gpu_ms(lambda m=m: (cert.setup_positions(xd, kind == "ee"), cert.certify(n, m)))
              for m in ("stream", "plain")}
        boxes = int(res["boxes"].sum())
        flops = boxes * (FLOP_BOX_VF if kind == "vf" else FLOP_BOX_EE)
        ref = [None]
        t_cpu = _cpu_s(lambda: ref.__setitem__(0, oracle.tuv_cert(kind == "ee", x, threads=1)))
        hit = res["hit"] != 0
        same = bool(np.array_equal(hit, ref[0]["hit"] != 0) and np.array_equal(res["toi"][hit], ref[0]["toi"][hit]))
        ach = flops / (ms["stream"] * 1e-3) / 1e12
        _row(f"tuv_certificate_{kind}", f"plain query set, {n} {kind.upper()} queries (tolerance 1e-6, ms 0)",
             {"value": n / (ms["stream"] * 1e-3), "unit": "queries/s", "ms": ms["stream"],
              "plain_form_ms": ms["plain"], "boxes": boxes, "autotuned": tuned},
             {"value": n / t_cpu, "unit": "queries/s", "cores": 1, "kind": "port",
              "sample": f"all {n} queries through the oracle's (t,u,v) restatement"},
             {"bound": "fp64-valu", "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
              "frac": ach / FP64_PEAK_TFLOPS, "traffic": None},
             {"identical_hits_and_toi": same})

    # reference interval certificate (CertificateEngine semantics)
    log("rows: reference interval certificate")
    cfg = nat.cert_config()
    for fam, kind, x in ((1, "point_triangle", qs.vf), (2, "edge_edge", qs.ee)):
        prims = nat.prims_from_positions(x)
        d_in = torch.from_numpy(prims.view(np.uint8).copy()).to(dev)
        d_out = torch.zeros(len(prims) * nat.PRIM_RESULT_DTYPE.itemsize, dtype=torch.uint8, device=dev)

        def run():
            nat.call("p2c_cert_ref_batch", fam, d_in.data_ptr(), len(prims), 0.0, 1.0, nat.ptr(cfg),
                     d_out.data_ptr(), None, nat.stream_handle())
        run()
        torch.cuda.synchronize()
        ms = _gpu_ms(run)
        got = d_out.cpu().numpy().view(nat.PRIM_RESULT_DTYPE)
        compiled = oracle.ref_available()
        f = oracle.ref_cert_compiled if compiled else oracle.ref_cert
        ref = [None]
        t_cpu = _cpu_s(lambda: ref.__setitem__(0, f(fam, oracle.prims_from_positions(x))))
        # the compiled reference does not count subdivision nodes; every other common field must match
        fields = [k for k in nat.PRIM_RESULT_DTYPE.names if k in ref[0].dtype.names and not (compiled and k == "nodes")]
        same = all(np.array_equal(got[k].astype(np.float64), ref[0][k].astype(np.float64)) for k in fields)
        _row(f"reference_certificate_{kind}", f"plain query set, {len(prims)} primitives, eps_time 1e-4, depth 32",
             {"value": len(prims) / (ms * 1e-3), "unit": "primitives/s", "ms": ms,
              "nodes": int(got["nodes"].sum())},
             {"value": len(prims) / t_cpu, "unit": "primitives/s", "cores": 1,
              "kind": "reference" if compiled else "port",
              "sample": "all primitives through " + ("the compiled certificate_engine.cpp" if compiled else
                                                     "the oracle restatement")},
             None, {"bit_exact": same})

    # exact work queue (ProcessExactWorkQueueCpu + refinement children) on the device
    log("rows: exact work queue")
    from p2cccd import exact_queue as eq

    qe, qpts, qees = workloads.exact_queue(100_000)
    dq = eq.DeviceQueue(qpts, qees, dev)
    d_entries = dq.upload_entries(qe)
    pt_pairs, ee_pairs = eq._pair_bounds(qe)
    refine = eq.RefineConfig()
    holder = {}
    ms = _gpu_ms(lambda: holder.__setitem__(0, dq.run(d_entries, pt_pairs, ee_pairs, eq.QueueConfig(), refine)))
    got_c = holder[0][0].cpu().numpy().view(eq.CERTIFICATE_DTYPE)
    got_r = holder[0][1].cpu().numpy().view(eq.AUDIT_ROW_DTYPE)
    compiled = oracle.ref_available()
    ref = [None]
    t_cpu = _cpu_s(lambda: ref.__setitem__(0, oracle.queue_process(qe, qpts, qees, compiled=compiled)))
    same = bool(np.array_equal(got_c.view(np.uint8), ref[0][0].view(np.uint8)) and
                np.array_equal(got_r.view(np.uint8), ref[0][1].view(np.uint8)))
    _row("exact_work_queue", f"{len(qe)} work items over 10k point-triangle and 10k edge-edge primitives (masks 1-3, "
         "0-4 primitives per family, eps_time 1e-4, depth 32): certificates, audit rows, refinement children",
         {"value": len(qe) / (ms * 1e-3), "unit": "work items/s", "ms": ms, "children": int(holder[0][3])},
         {"value": len(qe) / t_cpu, "unit": "work items/s", "cores": 1, "kind": "reference" if compiled else "port",
          "sample": "the whole queue through " + ("the compiled ProcessExactWorkQueueCpu" if compiled else
                                                  "the oracle restatement")},
         None, {"certificates_and_audit_rows_identical": same})

    # external-batch candidates (GenerateCandidatesForExternalBatch)
    log("rows: external batch candidates")
    import p2cccd_cpp as cpp
    from types import SimpleNamespace as NS

    n_ext = 100_000
    rng = np.random.default_rng(7)
    xq = np.zeros(n_ext, oracle.EXT_QUERY)
    xq["source_query_id"] = np.arange(n_ext) * 3 + 1
    xq["source_query_index"] = np.arange(n_ext)
    xq["family"] = rng.integers(0, 2, n_ext)
    xq["v0"] = rng.uniform(-1, 1, (n_ext, 1, 3)) + rng.uniform(-0.2, 0.2, (n_ext, 4, 3))
    xq["v1"] = xq["v0"] + rng.normal(0, 0.05, (n_ext, 4, 3))
    def ext_batch(rows):
        return NS(queries=[NS(query_id=int(r["source_query_id"]), source_query_index=int(r["source_query_index"]),
                              family="vf" if r["family"] == 0 else "ee", vertices_t0=r["v0"].tolist(),
                              vertices_t1=r["v1"].tolist(), ground_truth_collides=None, box_pair=None)
                           for r in rows])

    batch = ext_batch(xq)
    cpp.generate_candidates_for_external_batch(ext_batch(xq[:1000]))  # warm
    t0 = time.perf_counter()
    res = cpp.generate_candidates_for_external_batch(batch)
    wall = time.perf_counter() - t0
    # the reference tests every proxy pair of the batch (its cost grows with n^2): a 5,000-query sample, and the
    # device on the same sample for the parity check
    n_cpu = 5000
    cpu, same = None, None
    if oracle.ref_available():
        ref = [None]
        t_cpu = _cpu_s(lambda: ref.__setitem__(0, oracle.ref_external_batch(xq[:n_cpu])))
        small = cpp.generate_candidates_for_external_batch(ext_batch(xq[:n_cpu]))
        same = bool(np.array_equal(np.array([c.candidate_id for c in small.candidates], np.uint64),
                                   ref[0][0]["candidate_id"]))
        cpu = {"value": n_cpu / t_cpu, "unit": "queries/s", "cores": 1, "kind": "reference",
               "sample": f"the first {n_cpu} queries through the compiled external_batch_candidate.cpp (all proxy "
                         "pairs: per query its cost grows with the batch, so the 100k batch runs slower per query)"}
    _row("external_batch_candidates", f"{n_ext} external CCD queries (VF and EE, starts U[-1,1]^3, jitter 0.2, "
         f"motion N(0, 0.05)): {len(res.candidates)} candidates",
         {"value": n_ext / (res.timing.trace_ms * 1e-3), "unit": "queries/s", "device_ms": res.timing.trace_ms,
          "call_ms": wall * 1e3, "note": "call_ms adds the pybind conversion of the query objects and the copies"},
         cpu, None, {"identical_candidate_ids_on_the_sample": same})

    # full plain pipeline on two level-6 icospheres
    log("rows: two level-6 icospheres, full step")
    sc = scenes.two_icospheres()
    pipe = CcdPipeline(sc, device=dev)
    pipe.step_batched()
    lat = pipe.timed(pipe.step_batched, 10)
    rl, _, _ = stream_roofline(pipe, 5)
    pipe.step_plain()
    plain = pipe.timed(pipe.step_plain, 3)
    npairs = pipe.last.pairs
    cpu = cpu_baseline(sc, 100_000)
    _row("icospheres_step", f"{sc.name}: {npairs} candidates per step",
         {"value": npairs / (np.mean(lat) * 1e-3), "unit": "pairs/s", "ms_per_step": float(np.mean(lat)),
          "p99_ms": percentile(lat, 99), "plain_ms_per_step": float(np.mean(plain))}, cpu, rl,
         "covered by tests/test_pipeline_gpu.py (identical decisions and TOI against the oracle)")

    # candidate sets alone (swept boxes, LBVH build, count, fill) against the
    # reference's sort-and-sweep, restated by the oracle, on the same scene
    log("rows: candidate sets (icospheres)")
    from p2cccd.broadphase import BroadPhase

    bp = BroadPhase.for_scene(sc, dev)
    x0d, x1d = torch.from_numpy(sc.x0).to(dev), torch.from_numpy(sc.x1).to(dev)
    vf_d, ee_d = bp.candidates(x0d, x1d, sc.thickness)

    def sets():
        bp.build(x0d, x1d, sc.thickness)
        bp.count()
        bp.emit(vf_d, ee_d)
    ms_bp = _gpu_ms(sets, 5)
    t_c = time.perf_counter()
    ref_c = oracle.candidates(sc)
    t_sweep = time.perf_counter() - t_c
    same_sets = all(np.array_equal(np.unique(g.cpu().numpy(), axis=0), np.unique(ref_c[k], axis=0))
                    for k, g in (("vf", vf_d), ("ee", ee_d)))
    n_c = len(ref_c["vf"]) + len(ref_c["ee"])
    _row("candidate_sets_icospheres", f"{sc.name}: swept boxes, two LBVHs, {n_c} VF+EE candidates",
         {"value": n_c / (ms_bp * 1e-3), "unit": "candidates/s", "ms": ms_bp},
         {"value": n_c / t_sweep, "unit": "candidates/s", "cores": 1, "kind": "port",
          "sample": f"the oracle's sort-and-sweep over the whole scene ({t_sweep:.1f} s)"},
         None, {"identical_sets": same_sets})
    del bp

    # learned group scheduling: the two scorer kernels
    log("rows: scorer kernels")
    packed = stpf.load_stpf(stpf.DEFAULT_PRIORITY_WEIGHTS, dev)
    rng = np.random.default_rng(3)
    from p2cccd import train_stpf

    # the group features of the icosphere step's spatial cells (VF and EE), repeated to 65,536 groups: the
    # trained first layer carries the features' normalisation, so inputs follow their real distribution
    cells = train_stpf.scene_cells(TuvCertifier(device=dev), sc)
    real = torch.cat([d["gfeat"] for d in cells])
    feats = real.repeat((65536 + real.shape[0] - 1) // real.shape[0], 1)[:65536].contiguous()
    ms = _gpu_ms(lambda: stpf.stpf_priority(packed, feats))
    model = stpf.StpfPriority()
    p = packed.cpu()
    o = 0
    for prm in (model.l1.weight, model.l1.bias, model.l2.weight, model.l2.bias, model.head.weight, model.head.bias):
        prm.data.copy_(p[o:o + prm.numel()].view_as(prm))
        o += prm.numel()
    torch.set_num_threads(1)
    fc = feats.cpu()
    with torch.no_grad():
        want = model(fc)
        t_cpu = _cpu_s(lambda: model(fc))
    got = stpf.stpf_priority(packed, feats).cpu()
    ach = 65536 * STPF_FLOPS_ROW / (ms * 1e-3) / 1e12
    _row("stpf_priority_scorer", f"65,536 groups x 32 FP32 features ({real.shape[0]} spatial cells of the icosphere "
         "step, repeated), 32-128-128-1 GELU, sigmoid",
         {"value": 65536 / (ms * 1e-3), "unit": "groups/s", "ms": ms},
         {"value": 65536 / t_cpu, "unit": "groups/s", "cores": 1, "kind": "port",
          "sample": "the same model in torch on the CPU, FP32, one thread"},
         {"bound": "fp32", "achieved": ach, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
          "frac": ach / FP32_PEAK_TFLOPS, "traffic": None},
         {"max_abs_diff_vs_torch_fp32": float((got - want).abs().max())})
    bp = pipe.bp
    w = torch.from_numpy(rng.normal(0, 0.2, stpf.TINY_STPF_WEIGHTS)).to(dev)
    w[32:64] = 1.0  # std slots
    t = sc.topo
    x0, x1 = pipe.x0, pipe.x1
    vf = pipe.vf[: pipe.last.n_vf]
    n = int(vf.shape[0])
    topo = torch.from_numpy(np.ascontiguousarray(t.faces)).to(dev)
    oa = torch.from_numpy(np.ascontiguousarray(t.vertex_object)).to(dev)
    ob = torch.from_numpy(np.ascontiguousarray(t.face_object)).to(dev)
    ms = _gpu_ms(lambda: stpf.tiny_scores("vf", x0, x1, topo, vf, n, oa, ob, "learned", w))
    sample = rng.choice(n, size=min(n, 200_000), replace=False)
    pv = vf.cpu().numpy()[sample]
    vid = np.concatenate([pv[:, :1], t.faces[pv[:, 1]]], axis=1)
    rows = np.concatenate([sc.x0[vid], sc.x1[vid]], axis=1)
    keys = np.zeros(len(sample), np.uint64)
    t_cpu = _cpu_s(lambda: oracle.tiny_scores(0, True, rows, sample.astype(np.uint64), keys, w.cpu().numpy()))
    ach = n * TINY_FLOPS_ROW / (ms * 1e-3) / 1e12
    _row("tiny_stpf_scores", f"{n} VF candidates of the icosphere step, learned policy (FP64)",
         {"value": n / (ms * 1e-3), "unit": "candidates/s", "ms": ms},
         {"value": len(sample) / t_cpu, "unit": "candidates/s", "cores": 1, "kind": "port",
          "sample": f"{len(sample)} sampled candidates through the oracle's restatement"},
         {"bound": "fp64-valu", "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
          "frac": ach / FP64_PEAK_TFLOPS, "traffic": None},
         "covered by tests/test_scorer_gpu.py")
    del pipe

    # proxy-scene raw candidates + compaction
    log("rows: proxy scene")
    s = oracle.random_proxy_scene(np.random.default_rng(31), 20_000, 8, 4, extent=1.0, size=0.08)
    arr = proxy._device_arrays(s, dev)
    holder = {}

    def gen():
        raw = proxy.raw_hits(91, arr)
        holder["raw"] = raw
        holder["cand"] = proxy.compact(91, arr, raw)
    gen()
    ms = _gpu_ms(gen, 3)
    # the count pass alone, for its roofline: 8 FP64 compares per pair test (two
    # time, six box), at the FP64 VALU op rate (half the FMA-counted peak)
    n20 = arr["t0"].numel()
    ps = proxy._set(91, arr)
    counts = torch.zeros(n20, dtype=torch.int32, device=dev)
    wsb = int(nat.lib.p2c_proxy_raw_workspace_bytes(n20))
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    ms_count = _gpu_ms(lambda: nat.call("p2c_proxy_raw_count", ctypes.addressof(ps), counts.data_ptr(),
                                        ws.data_ptr(), wsb, nat.stream_handle()), 5)
    cmp_rate = (n20 * (n20 - 1) // 2) * 8 / (ms_count * 1e-3) / 1e12
    compiled = oracle.ref_available()
    f = oracle.ref_proxy_generate if compiled else oracle.proxy_generate
    ref = [None]
    t_cpu = _cpu_s(lambda: ref.__setitem__(0, f(91, s)))
    r_raw, r_cand = ref[0][0], ref[0][1]
    same = len(r_raw) == holder["raw"].numel() // 64 and len(r_cand) == len(holder["cand"]) and bool(
        np.array_equal(holder["cand"]["candidate_id"], r_cand["candidate_id"].astype(np.uint64)))
    npair = 20_000 * 19_999 // 2
    _row("proxy_scene_candidates", f"20,000 proxies, 8 objects, 4 slabs: {len(r_raw)} raw hits, "
         f"{len(r_cand)} candidates",
         {"value": npair / (ms * 1e-3), "unit": "pair tests/s", "ms": ms, "count_pass_ms": ms_count},
         {"value": npair / t_cpu, "unit": "pair tests/s", "cores": 1, "kind": "reference" if compiled else "port",
          "sample": "the whole scene through " + ("the compiled candidate_buffer.cpp" if compiled else "the oracle")},
         {"bound": "fp64-valu", "kernel": "proxy_rows_kernel (count pass)", "achieved": cmp_rate,
          "peak": FP64_PEAK_TFLOPS / 2, "unit": "T compares/s", "frac": cmp_rate / (FP64_PEAK_TFLOPS / 2),
          "traffic": None},
         {"identical_counts_and_ids": same})

    # serving: Poisson requests over 256 resident scenes, HBM budget below their footprint
    log("rows: serving")
    table, reqs = serve.poisson_workload(256, 600, 150.0)
    fp = serve.footprint(table)
    srv = serve.Server(table, int(0.65 * fp))
    warm = [serve.step_request(-1 - i, sid, table[sid], 99 + i) for i, sid in enumerate(sorted(table))]
    srv.run(warm, realtime=False)
    rep = serve.serve_timed(srv, reqs)
    sat = serve.serve_timed(srv, serve.poisson_requests(table, 4000, 4000.0, seed=77))
    _row("serving", f"256 resident scenes (1k-500k triangles, {fp / 2**30:.2f} GiB resident), HBM budget 65% of "
         "that; 600 Poisson requests at 150 Hz, then 4,000 at 4,000 Hz (saturated)",
         {"value": rep["pairs_per_s"], "unit": "pairs/s", "p50_ms": rep["p50_ms"], "p99_ms": rep["p99_ms"],
          "requests": rep["requests"], "evictions": rep["evictions"], "scene_loads": rep["scene_loads_during_run"],
          "saturated_pairs_per_s": sat["pairs_per_s"], "saturated_p50_ms": sat["p50_ms"],
          "saturated_p99_ms": sat["p99_ms"], "saturated_evictions": sat["evictions"],
          "saturated_batches": sat["batches"]}, None, None,
         "covered by tests/test_serve_gpu.py and tests/test_configs_gpu.py (each result identical alone and in the "
         "mix; 32 scenes under evictions against the oracle)")
    return 0


def served_form(scene, requests: int, want: tuple, cfg=None) -> dict:
    """The same step served through serve.Server (continuous batching over
    two lanes): requests alternate between two resident copies of the scene,
    so one lane's broad phase overlaps the other's certificate.  First all
    queued at once (saturated throughput), then arriving as a Poisson stream
    at 80% of that rate (latency).  want: (n_vf, n_ee, hits) of the batched
    pipeline on the same step, which every served request must reproduce;
    cfg: the certificate parameters (the batched step's tuned ones)."""
    import numpy as np

    from p2cccd import serve

    budget = 4 * serve.ResidentScene.estimate_bytes(scene)
    srv = serve.Server({0: scene, 1: scene}, budget, cfg=cfg)

    def req(k, t=0.0):
        return serve.Request(k, k % 2, scene.x0, scene.x1, t)

    srv.run([req(-1 - k) for k in range(4)], realtime=False)  # scenes resident, lanes sized
    t0 = time.perf_counter()
    done = srv.run([req(k) for k in range(requests)], realtime=False)
    sat = serve.report(done, time.perf_counter() - t0)
    rate = 0.8 * sat["pairs_per_s"] / max(want[0] + want[1], 1)
    arrivals = np.cumsum(np.random.default_rng(1).exponential(1.0 / rate, requests))
    t0 = time.perf_counter()
    done2 = srv.run([req(k, float(a)) for k, a in enumerate(arrivals)], realtime=True)
    lat = serve.report(done2, time.perf_counter() - t0)
    # unloaded: one request at a time
    done3 = [srv.run([req(k)], realtime=False)[0] for k in range(requests)]
    alone = [c.latency_ms for c in done3]
    same = all((c.n_vf, c.n_ee, len(c.vf_hits) + len(c.ee_hits)) == want for c in done + done2 + done3)
    return {"pairs_per_s_saturated": sat["pairs_per_s"], "p50_ms_saturated": sat["p50_ms"],
            "requests": requests, "rate_hz": rate, "pairs_per_s": lat["pairs_per_s"], "p50_ms": lat["p50_ms"],
            "p99_ms": lat["p99_ms"], "mean_ms": lat["mean_ms"], "p50_ms_unloaded": percentile(alone, 50),
            "p99_ms_unloaded": percentile(alone, 99), "identical_to_batched": same,
            "form": "serve.Server, two resident copies of the scene, requests alternating; saturated (all queued), "
                    "then Poisson arrivals at 80% of the saturated rate, then one request at a time (unloaded)"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--workload", choices=["cloth", "icospheres"], default="cloth")
    ap.add_argument("--cpu-sample", type=int, default=100_000, help="candidates certified for the CPU baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default certificate parameters")
    ap.add_argument("--probe", type=int, default=0,
                    help="profiling probe: run this many batched steps outside the graph, then a 1 GiB device copy "
                         "(the calibration kernel for HBM byte counters); prints the stream-kernel times")
    ap.add_argument("--rows", action="store_true",
                    help="per-row suite: each hot-path entry point on its workload beside the reference CPU path "
                         "(one JSON line per row)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from p2cccd import scenes
    from p2cccd.pipeline import CcdPipeline

    if args.rows:
        return rows_suite()

    dev = torch.device("cuda:0")
    log("building scene")
    scene = scenes.cloth_on_sphere() if args.workload == "cloth" else scenes.two_icospheres()
    pipe = CcdPipeline(scene, device=dev)

    if args.probe > 0:
        return probe(pipe, args.probe)

    log("warm-up")
    for _ in range(max(args.warmup, 1)):
        pipe.step_batched()
    torch.cuda.synchronize(dev)
    tuned = None
    if not args.no_autotune:
        log("autotune (stream-kernel launch and queue parameters, exact answers required)")
        tuned = pipe.autotune()
        log(f"autotune: {tuned}")
        pipe.step_batched()  # recapture the graph with the tuned parameters
        torch.cuda.synchronize(dev)

    # timed region: K graph-replayed batched steps, each synchronised
    lat = []
    pairs = 0
    torch.cuda.synchronize(dev)
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        c = pipe.step_batched()
        lat.append((time.perf_counter() - t0) * 1e3)
        pairs += c.pairs
    torch.cuda.synchronize(dev)
    total_s = time.perf_counter() - t_start
    r = pipe.results()
    log(f"batched: {pairs / total_s:.3e} pairs/s, {total_s / args.steps * 1e3:.2f} ms/step")

    # dominant kernel: the stream certificate kernel, timed with HIP events on
    # its launch stream (the current torch stream), then the wave pass
    rl, stage_ms, hc = stream_roofline(pipe, args.steps)
    log(f"stream kernel {rl['avg_ms']:.2f} ms ({rl['boxes_per_launch']} boxes), certificate stage "
        f"{np.mean(stage_ms):.2f} ms, wave pass took {hc} pairs")

    bp_rl = broadphase_roofline(pipe, args.steps)
    log(f"broad phase {bp_rl['avg_ms']:.2f} ms, {bp_rl['traffic'] / 1e9:.2f} GB algorithmic, "
        f"{bp_rl['achieved']:.0f} GB/s")

    # plain one-launch-per-stage form, same scene
    log("plain form")
    pipe.step_plain()
    plain_lat = pipe.timed(pipe.step_plain, args.plain_steps)
    rp = pipe.results()
    plain_pairs_s = rp["vf"].shape[0] + rp["ee"].shape[0]
    plain_rate = plain_pairs_s / (np.mean(plain_lat) * 1e-3)

    hits = int(r["vf_res"]["hit"].sum() + r["ee_res"]["hit"].sum())
    log("served form")
    served = served_form(scene, 2 * args.steps, (pipe.last.n_vf, pipe.last.n_ee, hits),
                         dataclasses.replace(pipe.cert.cfg))
    served["speedup_vs_plain"] = served["pairs_per_s_saturated"] / plain_rate
    log(f"served: {served['pairs_per_s_saturated']:.3e} pairs/s saturated; at {served['rate_hz']:.1f} Hz "
        f"p50 {served['p50_ms']:.2f} ms p99 {served['p99_ms']:.2f} ms; identical {served['identical_to_batched']}")
    line = {
        "metric": METRIC,
        "value": pairs / total_s,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_s / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": scene.name + (" (256x256 folded cloth self-contact on a level-5 sphere)"
                                      if args.workload == "cloth" else " (two level-6 icospheres)"),
            "pairs_per_step": pipe.last.pairs,
            "vf_pairs": pipe.last.n_vf,
            "ee_pairs": pipe.last.n_ee,
            "certificate": "tuv inclusion, tolerance 1e-6, ms 0, t_max 1, max_itr 1e6",
            "execution": "batched: device counts + stream certificate + wave pass + HIP graph replay",
            "autotuned": tuned,
        },
        "p50_ms": percentile(lat, 50),
        "p99_ms": percentile(lat, 99),
        "hits": hits,
        "replayed": int(r["replayed"]),
        "plain_pairs_per_s": plain_rate,
        "plain_ms_per_step": float(np.mean(plain_lat)),
        "speedup_vs_plain": (pairs / total_s) / plain_rate,
        "certificate_ms": float(np.mean(stage_ms)),
        "wave_pass_pairs": hc,
        "roofline": dict(rl, traffic=profiled_traffic("tuv_stream_kernel") if args.workload == "cloth" else None,
                         traffic_source="rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE "
                                        "(profiles/hbm_traffic_cloth256_step.md)"),
        "broad_phase_roofline": bp_rl,
        "served": served,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, 1 thread, then every host core)")
        line["cpu_baseline"] = cpu_baseline(scene, args.cpu_sample)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
