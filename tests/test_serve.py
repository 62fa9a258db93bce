"""Serving admission (host logic): scenes stay within the HBM budget, the
least recently used idle scene is evicted first, scenes with a request in
flight are never evicted, and a scene larger than the budget is an error."""
import pytest

from p2cccd.serve import Residency


def test_lru_eviction_respects_budget_and_busy_scenes():
    r = Residency(100)
    for sid, b in ((1, 40), (2, 30), (3, 20)):
        assert r.plan(sid, b) == []
        r.add(sid, b)
    r.touch(1)                      # order now 2, 3, 1
    r.busy.add(2)
    assert r.plan(4, 30) == [3]     # 2 is busy, 3 is the oldest idle one
    assert r.plan(5, 100) is None   # would need the busy scene's bytes
    r.busy.clear()
    assert r.plan(5, 100) == [2, 3, 1]
    assert r.plan(1, 999) == []     # already resident


def test_scene_larger_than_budget_is_rejected():
    with pytest.raises(RuntimeError, match="more than the HBM budget"):
        Residency(10).plan(1, 11)
