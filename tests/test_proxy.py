"""Proxy-scene validation (host rules of src/cpp/rt_candidate/proxy_scene.cpp:
18-154) with the reference's messages, and the density arithmetic against the
oracle on the golden scenes.  No kernel is launched."""
import numpy as np
import pytest

import oracle
from p2cccd import proxy

from test_oracle_proxy import golden_scenes


def _scene():
    return oracle.random_proxy_scene(np.random.default_rng(1), 20, 2, 2)


@pytest.mark.parametrize("mutate,msg", [
    (lambda s: None, None),
    (lambda s: s["proxy_type"].__setitem__(0, 3), "unsupported proxy type"),
    (lambda s: s["t1"].__setitem__(0, s["t0"][0]), "finite t0 < t1"),
    (lambda s: s["bounds"].__setitem__((0, 0), np.nan), "non-finite value"),
    (lambda s: s["bounds"].__setitem__((0, 0), s["bounds"][0, 3] + 1), "min cannot exceed max"),
    (lambda s: s["t1"].__setitem__(np.flatnonzero(s["slab_id"] == s["slab_id"][0])[1], 0.75),
     "one global time interval"),
])
def test_validation_messages(mutate, msg):
    s = _scene()
    mutate(s)
    if msg is None:
        proxy.validate_proxy_scene(5, s)
    else:
        with pytest.raises(ValueError, match=msg):
            proxy.validate_proxy_scene(5, s)
    with pytest.raises(ValueError, match="query_id is required"):
        proxy.validate_proxy_scene(0, _scene())


def test_overlapping_slabs_are_rejected():
    s = _scene()
    s["slab_id"][:] = np.arange(20) % 2
    s["t0"][:] = np.where(s["slab_id"] == 0, 0.0, 0.25)
    s["t1"][:] = np.where(s["slab_id"] == 0, 0.5, 1.0)
    with pytest.raises(ValueError, match="must not overlap"):
        proxy.validate_proxy_scene(5, s)


def test_density_arithmetic_matches_oracle():
    for q, scene, raw, cand, dens in golden_scenes():
        c = np.zeros(len(cand), proxy.CANDIDATE_DTYPE)
        c["rt_hit_count"] = cand["rt_hit_count"]
        d = proxy.density(scene, len(raw), c)
        for f in oracle.DENSITY.names:
            assert d[f] == dens[f], f
