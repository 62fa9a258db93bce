"""Learned group scheduling end to end on a held-out two-icosphere scene
(level 5, a seed outside the training variants): the trained STPF priority
(p2cccd/data/stpf_priority.npz, produced by python -m p2cccd.train_stpf)
orders spatial cells; exact certificates decide every cell in both the
scheduled and the plain order, the decisions are identical and equal the
cells holding a certified contact, and the scheduled order reaches its
detections with fewer certificates on average.  (The total over all cells
depends only on the order inside each cell and is not asserted: the
proximity order is a heuristic there.)"""
import pytest
import torch

from p2cccd import scenes, stpf, train_stpf

pytestmark = pytest.mark.gpu


def test_scheduled_order_detects_earlier_with_identical_decisions(gpu):
    packed = stpf.load_stpf(stpf.DEFAULT_PRIORITY_WEIGHTS, torch.device("cuda"))
    rep = train_stpf.evaluate(packed, scenes.two_icospheres(level=5, gap=0.01, seed=77))
    assert rep["kinds"]
    for kind, r in rep["kinds"].items():
        assert r["decisions_identical"]
        print(kind, r)
        if r["positive_cells"]:
            assert r["scheduled"]["mean_detection"] < r["plain"]["mean_detection"], kind
