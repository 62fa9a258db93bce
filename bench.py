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


def percentile(v, q):
    import numpy as np

    return float(np.percentile(np.asarray(v), q))


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
        t1 = time.perf_counter()
        oracle.tuv_cert(kind == "ee", x, threads=1)
        t_cert += (time.perf_counter() - t1) * len(pairs) / max(len(idx), 1)
        sampled += len(idx)
    step_s = t_sweep + t_cert
    return {
        "value": n_total / step_s,
        "unit": "pairs/s",
        "cores": 1,
        "kind": "port",
        "sample": f"full sort-and-sweep of the scene ({n_total} candidates, {t_sweep:.1f} s) + certificate on "
                  f"{sampled} uniformly sampled candidates scaled to all ({t_cert:.1f} s est.), 1 thread",
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
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from p2cccd import scenes
    from p2cccd.pipeline import CcdPipeline

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
    stream_ms, stage_ms = [], []
    for _ in range(args.steps):
        ev = tuple(torch.cuda.Event(enable_timing=True) for _ in range(3))
        pipe.step_batched_timed(ev)
        stream_ms.append(ev[0].elapsed_time(ev[2]))
        stage_ms.append(ev[0].elapsed_time(ev[1]))
    rt = pipe.results()
    n_vf = pipe.last.n_vf
    boxes = np.concatenate([rt["vf_res"]["boxes"], rt["ee_res"]["boxes"]]).astype(np.int64)
    # boxes the wave pass added for the pairs it took over (their saved walks
    # hold the count the stream kernel reached)
    cert = pipe.cert
    hc = cert.heavy_count()
    heavy = cert.heavy_list[:hc].long()
    saved = cert.states.view(-1, cert.state_bytes)[heavy][:, 32:36].contiguous().view(torch.int32)
    heavy = heavy.cpu().numpy()
    wave_part = np.maximum(boxes[heavy] - saved.cpu().numpy()[:, 0].astype(np.int64), 0)
    is_vf = np.arange(len(boxes)) < n_vf
    stream_vf = int(boxes[is_vf].sum() - wave_part[heavy < n_vf].sum())
    stream_ee = int(boxes[~is_vf].sum() - wave_part[heavy >= n_vf].sum())
    flops = stream_vf * FLOP_BOX_VF + stream_ee * FLOP_BOX_EE
    stream_avg_ms = float(np.mean(stream_ms))
    achieved = flops / (stream_avg_ms * 1e-3) / 1e12
    log(f"stream kernel {stream_avg_ms:.2f} ms ({stream_vf + stream_ee} boxes), certificate stage "
        f"{np.mean(stage_ms):.2f} ms, wave pass took {hc} pairs")

    # plain one-launch-per-stage form, same scene
    log("plain form")
    pipe.step_plain()
    plain_lat = pipe.timed(pipe.step_plain, args.plain_steps)
    rp = pipe.results()
    plain_pairs_s = rp["vf"].shape[0] + rp["ee"].shape[0]
    plain_rate = plain_pairs_s / (np.mean(plain_lat) * 1e-3)

    hits = int(r["vf_res"]["hit"].sum() + r["ee_res"]["hit"].sum())
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
        "roofline": {
            "bound": "fp64-valu",
            "kernel": "tuv_stream_kernel",
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": None,
            "avg_ms": stream_avg_ms,
            "flops_per_launch": flops,
            "boxes_per_launch": stream_vf + stream_ee,
        },
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, 1 thread)")
        line["cpu_baseline"] = cpu_baseline(scene, args.cpu_sample)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
