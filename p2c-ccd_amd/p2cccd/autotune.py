"""Autotuner for the batched certificate's launch and queue parameters.

The stream kernel's refill threshold, light budget, hand-over threshold and
resident blocks per CU, and the wave pass's grid, change only how work is
spread over the chip; the certificate's answers must not change.  The tuner
measures each candidate configuration on the caller's own pairs (HIP events
around the certificate launches, best of a few repeats) and admits it only if
every pair's hit flag and TOI are bit-identical to the plain one-lane-per-pair
kernel's.  Results are cached per device and workload signature in a JSON
file so a serving process tunes once.
"""
from __future__ import annotations

import dataclasses
import itertools
import json
import os
from dataclasses import dataclass
from pathlib import Path
from typing import Callable, Iterable

import numpy as np
import torch

from .tuv import TuvCertifier, TuvConfig

SPACE = {
    "refill_at": (4, 8, 16),
    "light_budget": (32, 512, 1024, 2048),
    "handover_at": (0, 8, 16),
    "blocks_per_cu": (2, 3, 4),
}
# tried after the grid, with the grid's winner: boxes a wave takes on a pair before it hands
# walks to waiting waves (0: the small-batch default, or never for a batch larger than the grid)
SPLIT_CHOICES = (0, 1024, 2048, 4096, 16384)


@dataclass
class Trial:
    params: dict
    ms: float
    correct: bool


def select(candidates: Iterable[dict], measure: Callable[[dict], float], check: Callable[[dict], bool]) -> tuple[dict,
                                                                                                                 list]:
    """The fastest candidate that passes `check`; every trial is returned."""
    trials = []
    best = None
    for p in candidates:
        ok = check(p)
        ms = measure(p) if ok else float("inf")
        trials.append(Trial(dict(p), ms, ok))
        if ok and (best is None or ms < best.ms):
            best = trials[-1]
    if best is None:
        raise RuntimeError("no candidate configuration passed the correctness gate")
    return best.params, trials


def grid(space: dict | None = None):
    space = space or SPACE
    keys = list(space)
    for vals in itertools.product(*(space[k] for k in keys)):
        yield dict(zip(keys, vals))


def _signature(n: int, boxes: int) -> str:
    name = torch.cuda.get_device_properties(0).name if torch.cuda.is_available() else "cpu"
    return f"{name}|pairs~{1 << max(n, 1).bit_length()}|boxes/pair~{round(boxes / max(n, 1), 0)}"


class CertificateTuner:
    """Tunes TuvConfig for the pairs currently set up in `cert`."""

    def __init__(self, cert: TuvCertifier, n: int, repeats: int = 3, cache: str | Path | None = None):
        self.cert, self.n, self.repeats = cert, n, repeats
        self.cache = Path(cache or os.environ.get("P2C_TUNE_CACHE", Path.home() / ".cache" / "p2c_tune.json"))
        cert.certify(n, "plain")
        torch.cuda.synchronize()
        r = self._snapshot()
        self.reference = r
        self.boxes = int(cert.result_tensor(n).view(torch.int64).view(n, 4)[:, 1].sum())

    def _snapshot(self):
        res = self.cert.result_tensor(self.n).view(torch.float64).view(self.n, 4)
        return res[:, 0].clone(), (res[:, 2].view(torch.int64) & 0xFFFFFFFF).clone()

    def _apply(self, p: dict) -> None:
        cfg = dataclasses.replace(self.cert.cfg, refill_at=p["refill_at"], light_budget=p["light_budget"],
                                  handover_at=p["handover_at"], blocks_per_cu=p["blocks_per_cu"],
                                  split_at=p.get("split_at", self.cert.cfg.split_at))
        self.cert.cfg = cfg
        props = torch.cuda.get_device_properties(self.cert.device)
        self.cert.grid_blocks = int(props.multi_processor_count) * cfg.blocks_per_cu

    def check(self, p: dict) -> bool:
        self._apply(p)
        self.cert.certify(self.n, "stream")
        torch.cuda.synchronize()
        toi, hit = self._snapshot()
        return bool(torch.equal(hit, self.reference[1]) and torch.equal(toi, self.reference[0]))

    def measure(self, p: dict) -> float:
        self._apply(p)
        best = float("inf")
        for _ in range(self.repeats):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            self.cert.certify(self.n, "stream", events=(e0, e1))
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1))
        return best

    def tune(self, space: dict | None = None, use_cache: bool = True) -> tuple[TuvConfig, list]:
        sig = _signature(self.n, self.boxes)
        cached = self._load().get(sig) if use_cache else None
        if cached is not None:
            self._apply(cached)
            if self.check(cached):
                return self.cert.cfg, []
        params, trials = select(grid(space), self.measure, self.check)
        params, more = select((dict(params, split_at=v) for v in SPLIT_CHOICES), self.measure, self.check)
        trials += more
        self._apply(params)
        db = self._load()
        db[sig] = params
        try:
            self.cache.parent.mkdir(parents=True, exist_ok=True)
            self.cache.write_text(json.dumps(db, indent=1))
        except OSError:
            pass
        return self.cert.cfg, trials

    def _load(self) -> dict:
        try:
            return json.loads(self.cache.read_text())
        except (OSError, ValueError):
            return {}


def summarize(trials: list) -> dict:
    ok = [t for t in trials if t.correct]
    ms = np.array([t.ms for t in ok])
    return {"tried": len(trials), "passed": len(ok), "best_ms": float(ms.min()) if ok else None,
            "median_ms": float(np.median(ms)) if ok else None, "worst_ms": float(ms.max()) if ok else None}
