"""Per-timestep CCD for one resident scene: LBVH broad phase, then the
(t,u,v) certificate on every VF and EE candidate.

Two execution forms with identical answers:

  plain    one launch per stage; the host reads the candidate counts back to
           size the pair buffers, and the certificate runs one lane per pair.
  batched  every stage enqueued back to back with the counts kept on the
           device, the certificate as the stream kernel (batched lane refill)
           plus the wave-cooperative pass, the whole step captured once in a
           HIP graph and replayed.

This is the reference's all-exact path (src/tools/scene_object_envelope_native
_ti_walltime.cpp:1239-1281: every candidate certified) on the device.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np
import torch

from .broadphase import BroadPhase
from .scenes import Scene
from .tuv import TUV_RESULT_DTYPE, TuvCertifier, TuvConfig, results_view


@dataclass
class StepCounts:
    n_vf: int
    n_ee: int

    @property
    def pairs(self) -> int:
        return self.n_vf + self.n_ee


class CcdPipeline:
    def __init__(self, scene: Scene, config: TuvConfig | None = None, device: str | torch.device = "cuda",
                 headroom: float = 1.25):
        self.scene = scene
        self.device = torch.device(device)
        self.bp = BroadPhase.for_scene(scene, self.device)
        self.x0 = torch.from_numpy(np.ascontiguousarray(scene.x0)).to(self.device)
        self.x1 = torch.from_numpy(np.ascontiguousarray(scene.x1)).to(self.device)
        self.thickness = scene.thickness
        self.cert = TuvCertifier(config, self.device)
        self.headroom = headroom
        self.batched_mode = "stream"
        # batched stream form: 64-byte pair records expanded from the scene's
        # vertex arrays (L2-resident) instead of 288-byte records
        self.compact_records = True
        self.cap_vf = self.cap_ee = 0
        self.vf = self.ee = None
        self.graph: torch.cuda.CUDAGraph | None = None
        self.last = StepCounts(0, 0)

    # -- buffers ------------------------------------------------------------
    def _ensure(self, n_vf: int, n_ee: int) -> bool:
        """Grows the pair buffers; returns True when they were reallocated."""
        if n_vf <= self.cap_vf and n_ee <= self.cap_ee and self.vf is not None:
            return False
        self.cap_vf = max(int(n_vf * self.headroom) + 1024, self.cap_vf)
        self.cap_ee = max(int(n_ee * self.headroom) + 1024, self.cap_ee)
        self.vf = torch.empty((self.cap_vf, 2), dtype=torch.int32, device=self.device)
        self.ee = torch.empty((self.cap_ee, 2), dtype=torch.int32, device=self.device)
        self.cert.reserve(self.cap_vf + self.cap_ee)
        self.graph = None
        return True

    def _read_counts(self) -> StepCounts:
        n_vf, n_ee, bad = (int(v) for v in self.bp.totals.cpu())
        if bad:
            raise RuntimeError(f"{bad} BVH traversals exceeded their step bound")
        return StepCounts(n_vf, n_ee)

    # -- plain: one launch per stage, host-sized buffers ---------------------
    def step_plain(self) -> StepCounts:
        bp, c = self.bp, self.cert
        bp.build(self.x0, self.x1, self.thickness)
        bp.count()
        counts = self._read_counts()  # host round trip to size the buffers
        self._ensure(counts.n_vf, counts.n_ee)
        bp.emit(self.vf, self.ee)
        c.setup_pairs(self.x0, self.x1, bp.faces, self.vf, False, counts.n_vf, out_offset=0)
        c.setup_pairs(self.x0, self.x1, bp.edges, self.ee, True, counts.n_ee, out_offset=counts.n_vf)
        c.certify(counts.pairs, "plain")
        torch.cuda.synchronize(self.device)
        self.last = counts
        return counts

    # -- batched: device-side counts, stream certificate, graph --------------
    def _enqueue_batched(self, events: tuple | None = None) -> None:
        bp, c = self.bp, self.cert
        bp.build(self.x0, self.x1, self.thickness)
        bp.count()
        bp.emit(self.vf, self.ee)
        lite = self.compact_records and self.batched_mode in ("stream", "plain")
        c.setup_pairs(self.x0, self.x1, bp.faces, self.vf, False, self.cap_vf, totals=bp.totals, lite=lite)
        c.setup_pairs(self.x0, self.x1, bp.edges, self.ee, True, self.cap_ee, totals=bp.totals, lite=lite)
        c.certify(self.cap_vf + self.cap_ee, self.batched_mode, totals=bp.totals, events=events)

    def _capture(self) -> None:
        s = torch.cuda.Stream(self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            self._enqueue_batched()  # warm-up outside the capture
        torch.cuda.current_stream(self.device).wait_stream(s)
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._enqueue_batched()
        self.graph = g

    def step_batched_timed(self, events: tuple) -> StepCounts:
        """One batched step outside the graph, with `events` recorded around
        the certificate kernels (same kernels as the graph; see
        TuvCertifier.certify)."""
        self._enqueue_batched(events)
        torch.cuda.synchronize(self.device)
        self.last = self._read_counts()
        return self.last

    def step_batched(self, use_graph: bool = True) -> StepCounts:
        if self.vf is None:
            self.bp.build(self.x0, self.x1, self.thickness)
            self.bp.count()
            c0 = self._read_counts()
            self._ensure(c0.n_vf, c0.n_ee)
        for _ in range(2):
            if use_graph:
                if self.graph is None:
                    self._capture()
                self.graph.replay()
            else:
                self._enqueue_batched()
            torch.cuda.synchronize(self.device)
            counts = self._read_counts()
            if not self._ensure(counts.n_vf, counts.n_ee):
                self.last = counts
                return counts
        raise RuntimeError("candidate counts kept outgrowing the pair buffers")

    # -- tuning --------------------------------------------------------------
    def autotune(self, space: dict | None = None, cache=None) -> dict:
        """Picks the certificate's launch and queue parameters on this scene's
        current pairs (p2cccd.autotune; every admitted configuration returns
        bit-identical results) and drops the captured graph."""
        from .autotune import CertificateTuner, summarize

        if self.vf is None:
            self.step_batched(use_graph=False)
        self._enqueue_batched()
        torch.cuda.synchronize(self.device)
        self.last = self._read_counts()
        tuner = CertificateTuner(self.cert, self.last.pairs, cache=cache)
        cfg, trials = tuner.tune(space)
        self.graph = None
        return {"refill_at": cfg.refill_at, "light_budget": cfg.light_budget, "handover_at": cfg.handover_at,
                "blocks_per_cu": cfg.blocks_per_cu, "split_at": cfg.split_at, **summarize(trials)}

    # -- results ------------------------------------------------------------
    def results(self) -> dict:
        """Host copies of the last step: candidate pairs and per-pair results
        (VF first, then EE, in emission order)."""
        n_vf, n_ee = self.last.n_vf, self.last.n_ee
        res = results_view(self.cert.result_tensor(n_vf + n_ee))
        return dict(vf=self.vf[:n_vf].cpu().numpy(), ee=self.ee[:n_ee].cpu().numpy(), vf_res=res[:n_vf],
                    ee_res=res[n_vf:n_vf + n_ee], replayed=self.cert.replay_count())

    def timed(self, fn, steps: int) -> list[float]:
        out = []
        for _ in range(steps):
            t0 = time.perf_counter()
            fn()
            out.append((time.perf_counter() - t0) * 1e3)
        return out
