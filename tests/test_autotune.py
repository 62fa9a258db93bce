"""The autotuner's selection rule (CPU): the fastest configuration that passes
the correctness gate wins; failing configurations are never chosen, however
fast; no passing configuration is an error."""
import pytest

from p2cccd.autotune import grid, select


def test_fastest_passing_configuration_wins():
    times = {1: 5.0, 2: 1.0, 3: 3.0, 4: 0.5}
    ok = {1: True, 2: True, 3: True, 4: False}  # 4 is fastest but changes answers
    params, trials = select(({"k": k} for k in times), lambda p: times[p["k"]], lambda p: ok[p["k"]])
    assert params == {"k": 2}
    assert [t.correct for t in trials] == [True, True, True, False]


def test_no_passing_configuration_raises():
    with pytest.raises(RuntimeError, match="correctness gate"):
        select([{"k": 1}], lambda p: 1.0, lambda p: False)


def test_grid_covers_the_space():
    space = {"a": (1, 2), "b": (3, 4, 5)}
    assert len(list(grid(space))) == 6
