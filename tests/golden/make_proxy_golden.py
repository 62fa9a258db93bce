"""Writes tests/golden/proxy_golden.npz from the reference's own compiled
proxy-scene candidate code (src/cpp/rt_candidate/candidate_buffer.cpp,
proxy_scene.cpp, candidate_stats.cpp built in place into oracle/_ref by
`make -C oracle ref`).

Run in the container that holds the reference:  python tests/golden/make_proxy_golden.py
The fixture holds inputs and expected outputs only.  Scenes: the scene of the
reference's candidate KAT (src/tests/cpp/test_rt_candidate.cpp:60-106, built
by the reference's BuildProxyScene) and seeded random scenes of 2-600
proxies with 1-5 objects and 1-4 slabs, some with duplicate compaction keys.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402

OUT = Path(__file__).resolve().parent / "proxy_golden.npz"


def main() -> None:
    rng = np.random.default_rng(20261203)
    scenes = [oracle.ref_proxy_kat_scene()]
    for k, (n, objs, slabs) in enumerate([(2, 2, 1), (30, 2, 1), (80, 3, 2), (200, 5, 4), (600, 4, 3), (150, 2, 1)]):
        s = oracle.random_proxy_scene(rng, n, objs, slabs, extent=0.5, size=0.5 if k == 5 else 0.25)
        if k == 5:
            s["patch_id"][:] = 0  # few distinct (object, patch) keys: many duplicates to compact
        scenes.append((1000 + k, s))
    data = {}
    for i, (q, s) in enumerate(scenes):
        raw, cand, dens = oracle.ref_proxy_generate(q, s)
        data[f"s{i}_query"] = np.array(q, np.uint64)
        for key, v in s.items():
            data[f"s{i}_{key}"] = v
        data[f"s{i}_raw"] = raw
        data[f"s{i}_cand"] = cand
        data[f"s{i}_density"] = np.array([dens])
        print(i, len(s["t0"]), "proxies", len(raw), "raw", len(cand), "candidates")
    data["count"] = np.array(len(scenes))
    np.savez_compressed(OUT, **data)


if __name__ == "__main__":
    main()
