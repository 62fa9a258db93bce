"""Autotuning on the GPU: the tuned certificate configuration gives the same
per-pair answers as the plain kernel, the choice is cached per device and
workload, and the pipeline keeps matching the oracle after tuning."""
import pytest
import torch

from p2cccd import scenes
from p2cccd.autotune import SPLIT_CHOICES
from p2cccd.pipeline import CcdPipeline

pytestmark = pytest.mark.gpu

SPACE = {"refill_at": (1, 8), "light_budget": (4, 1024), "handover_at": (0, 64), "blocks_per_cu": (1, 3)}


def test_tuned_configuration_is_exact_and_cached(gpu, tmp_path):
    s = scenes.cloth_on_sphere(n=48, sphere_level=3, cloth_size=1.5)
    p = CcdPipeline(s)
    p.step_plain()
    plain = p.results()
    info = p.autotune(SPACE, cache=tmp_path / "tune.json")
    assert info["passed"] == info["tried"] == 16 + len(SPLIT_CHOICES)  # the grid, then the split threshold
    assert (tmp_path / "tune.json").exists()
    p.step_batched()
    r = p.results()
    for kind in ("vf", "ee"):
        assert torch.equal(torch.from_numpy(r[kind]), torch.from_numpy(plain[kind]))
        assert (r[kind + "_res"]["hit"] == plain[kind + "_res"]["hit"]).all()
        assert (r[kind + "_res"]["toi"] == plain[kind + "_res"]["toi"]).all()
    again = p.autotune(SPACE, cache=tmp_path / "tune.json")
    assert again["tried"] == 0  # served from the cache after re-checking
