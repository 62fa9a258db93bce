"""gfx950 (t,u,v) inclusion certificate against the canonical oracle
(oracle/tuv_cert.c): identical hit decisions, bit-identical TOI, on the plain
20k-query workload, through both execution forms, and with a small work cap
that forces the fallback-replay queue (which must then reproduce the
canonical capped answer)."""
import numpy as np
import pytest
import torch

import oracle
from p2cccd import workloads
from p2cccd.tuv import FLAG_REPLAYED, TuvCertifier, TuvConfig

pytestmark = pytest.mark.gpu


def _check(got, ref, label):
    assert np.array_equal(got["hit"], ref["hit"]), label
    assert np.array_equal(got["toi"], ref["toi"]), label


@pytest.mark.parametrize("mode", ["plain", "stream", "persistent", "rounds"])
def test_plain_workload_matches_oracle(gpu, mode):
    qs = workloads.make_query_set(10_000, 10_000)
    cert = TuvCertifier()
    for is_ee, x in ((False, qs.vf), (True, qs.ee)):
        got = cert.certify_positions(x, is_ee, mode)
        ref = oracle.tuv_cert(is_ee, x, threads=8)
        _check(got, ref, (mode, is_ee))
        assert cert.replay_count() == 0
        assert got["hit"][(qs.vf_kind if not is_ee else qs.ee_kind) == 1].all()  # touching pairs all hit


@pytest.mark.parametrize("mode", ["plain", "stream", "persistent", "rounds"])
def test_forced_replay_reproduces_capped_oracle(gpu, mode):
    qs = workloads.make_query_set(2000, 2000, seed=31)
    cfg = TuvConfig(max_itr=40, heap_cap=1 << 12, replay_slots=16)
    cert = TuvCertifier(cfg)
    for is_ee, x in ((False, qs.vf), (True, qs.ee)):
        got = cert.certify_positions(x, is_ee, mode)
        ref = oracle.tuv_cert(is_ee, x, max_itr=40, heap_cap=1 << 12, threads=8)
        _check(got, ref, (mode, is_ee))
        replayed = (got["flags"] & FLAG_REPLAYED) != 0
        assert replayed.sum() == cert.replay_count() > 0
        assert np.array_equal(got["boxes"][replayed], ref["boxes"][replayed])


def test_tolerance_and_separation_variants(gpu):
    qs = workloads.make_query_set(1500, 1500, seed=44)
    for kw in (dict(tolerance=1e-4), dict(tolerance=1e-8), dict(ms=1e-3)):
        cert = TuvCertifier(TuvConfig(tolerance=kw.get("tolerance", 1e-6), ms=kw.get("ms", 0.0)))
        for is_ee, x in ((False, qs.vf), (True, qs.ee)):
            got = cert.certify_positions(x, is_ee, "persistent")
            ref = oracle.tuv_cert(is_ee, x, ms=kw.get("ms", 0.0), tolerance=kw.get("tolerance", 1e-6), threads=8)
            _check(got, ref, kw)


@pytest.mark.parametrize("light_budget", [1, 8, 64])
def test_wave_cooperative_pass_matches_oracle(gpu, light_budget):
    # A tiny lane budget sends most pairs through the wave-shared pass, whose
    # box order differs from every other path; decisions and TOI must not.
    qs = workloads.make_query_set(3000, 3000, seed=52)
    cert = TuvCertifier(TuvConfig(light_budget=light_budget))
    for is_ee, x in ((False, qs.vf), (True, qs.ee)):
        got = cert.certify_positions(x, is_ee, "persistent")
        ref = oracle.tuv_cert(is_ee, x, threads=8)
        _check(got, ref, (light_budget, is_ee))
        assert cert.heavy_count() > 0 and cert.replay_count() == 0


@pytest.mark.parametrize("budgets", [(1,), (1, 2), (2, 4, 8, 16)])
def test_rounds_with_tiny_budgets_match_oracle(gpu, budgets):
    # Walks are cut after a step or two and resumed from their saved state in
    # the next round; survivors of the last round restart in the wave pass.
    qs = workloads.make_query_set(3000, 3000, seed=61)
    cert = TuvCertifier(TuvConfig(round_budgets=budgets))
    for is_ee, x in ((False, qs.vf), (True, qs.ee)):
        got = cert.certify_positions(x, is_ee, "rounds")
        ref = oracle.tuv_cert(is_ee, x, threads=8)
        _check(got, ref, (budgets, is_ee))
        assert cert.heavy_count() > 0


@pytest.mark.parametrize("refill_at,light_budget,handover_at",
                         [(1, 4, 0), (8, 4, 8), (64, 4, 64), (1, 1024, 64), (8, 1024, 8), (64, 1024, 0)])
def test_stream_refill_and_handover_match_oracle(gpu, refill_at, light_budget, handover_at):
    # refill_at 1 refills a lane as soon as it idles, 64 only when the whole
    # wave is idle; light budget 4 sends most pairs, mid-walk, to the wave
    # pass; handover_at 64 hands every walk still open when the pairs run
    # out (some between two boxes) to the wave pass.
    qs = workloads.make_query_set(3000, 3000, seed=67)
    cert = TuvCertifier(TuvConfig(light_budget=light_budget, refill_at=refill_at, handover_at=handover_at,
                                  small_batch_budget=light_budget))
    for is_ee, x in ((False, qs.vf), (True, qs.ee)):
        got = cert.certify_positions(x, is_ee, "stream")
        ref = oracle.tuv_cert(is_ee, x, threads=8)
        _check(got, ref, (refill_at, light_budget, handover_at, is_ee))
        assert cert.replay_count() == 0
        if light_budget == 4 or handover_at == 64:
            assert cert.heavy_count() > 0


def test_single_lane_forms_walk_the_same_boxes(gpu):
    # Without a hand-over to the wave pass, the stream and rounds forms walk
    # each pair's boxes in the plain kernel's depth-first order, so even the
    # per-pair box counts agree.
    qs = workloads.make_query_set(4000, 4000, seed=73)
    for is_ee, x in ((False, qs.vf), (True, qs.ee)):
        plain = TuvCertifier().certify_positions(x, is_ee, "plain")
        big = TuvConfig(light_budget=1_000_000, small_batch_budget=1_000_000, round_budgets=(1_000_000,),
                        handover_at=0)
        for mode in ("stream", "rounds"):
            c = TuvCertifier(big)
            got = c.certify_positions(x, is_ee, mode)
            assert c.heavy_count() == 0
            _check(got, plain, (mode, is_ee))
            assert np.array_equal(got["boxes"], plain["boxes"]), (mode, is_ee)


@pytest.mark.parametrize("split_at,tasks", [(32, 1 << 16), (32, 300), (512, 1 << 16)])
def test_split_wave_pass_matches_oracle(gpu, split_at, tasks):
    # Pairs over split_at boxes in the wave pass are cut into pieces that
    # many waves certify with a shared best; with room for only 300 pieces
    # most splits are refused and those pairs finish in their own wave.
    qs = workloads.make_query_set(3000, 3000, seed=53)
    cert = TuvCertifier(TuvConfig(light_budget=8, split_at=split_at, split_tasks=tasks))
    for is_ee, x in ((False, qs.vf), (True, qs.ee)):
        got = cert.certify_positions(x, is_ee, "stream")
        ref = oracle.tuv_cert(is_ee, x, threads=8)
        _check(got, ref, (split_at, tasks, is_ee))
        pieces, split = cert.split_counts()
        assert split > 0 and 0 < pieces <= tasks and cert.replay_count() == 0


def test_split_wave_pass_deep_pairs_match_oracle(gpu):
    # Every third pair scaled by 1e4 splits some dimension more than 31
    # times: such pairs take the 64-bit-key wave walk and its split, their
    # pieces sharing the task queue with the jump walk's; two per family
    # run over max_itr and go to the canonical replay.
    qs = workloads.make_query_set(1500, 1500, seed=59)
    cfg = TuvConfig(light_budget=8, split_at=32, max_itr=20000, heap_cap=1 << 16)
    cert = TuvCertifier(cfg)
    for is_ee, x in ((False, qs.vf.copy()), (True, qs.ee.copy())):
        x[::3] *= 1e4
        got = cert.certify_positions(x, is_ee, "stream")
        ref = oracle.tuv_cert(is_ee, x, max_itr=20000, heap_cap=1 << 16, threads=8)
        _check(got, ref, is_ee)
        assert cert.split_counts()[1] > 0 and cert.replay_count() > 0


def test_split_wave_pass_forced_replay(gpu):
    # Split pairs whose pieces run out of budget (or whose total does) go to
    # the canonical replay like any other over-budget pair (about 150 pairs
    # per family exceed 500 boxes).
    qs = workloads.make_query_set(2000, 2000, seed=31)
    cfg = TuvConfig(max_itr=500, heap_cap=1 << 12, replay_slots=16, light_budget=4, split_at=8)
    cert = TuvCertifier(cfg)
    for is_ee, x in ((False, qs.vf), (True, qs.ee)):
        got = cert.certify_positions(x, is_ee, "stream")
        ref = oracle.tuv_cert(is_ee, x, max_itr=500, heap_cap=1 << 12, threads=8)
        _check(got, ref, is_ee)
        replayed = (got["flags"] & FLAG_REPLAYED) != 0
        assert replayed.sum() == cert.replay_count() > 0
        assert cert.split_counts()[1] > 0


def test_small_batch_budget_matches_oracle(gpu):
    # A batch with fewer pairs than the grid has lanes walks at most
    # small_batch_budget boxes per lane before the wave pass takes over.
    qs = workloads.make_query_set(3000, 3000, seed=71)
    for budget in (1, 32):
        cert = TuvCertifier(TuvConfig(small_batch_budget=budget))
        for is_ee, x in ((False, qs.vf), (True, qs.ee)):
            got = cert.certify_positions(x, is_ee, "stream")
            _check(got, oracle.tuv_cert(is_ee, x, threads=8), (budget, is_ee))
            assert cert.heavy_count() > 0 and cert.replay_count() == 0


@pytest.mark.parametrize("rule", ["python", "cpp"])
def test_tight_inclusion_csv_queries_on_device(gpu, rule):
    """Queries read from the hand-written Tight-Inclusion CSV fixtures
    (p2cccd.datasets) through the device certificate: identical to the
    oracle and conservative against the files' ground-truth labels (every
    labelled collision is reported)."""
    from pathlib import Path

    from p2cccd.datasets import load_queries

    root = Path(__file__).resolve().parent / "golden" / "tight_inclusion" / "handwritten"
    cert = TuvCertifier()
    for kind, is_ee in (("vertex-face", False), ("edge-edge", True)):
        x, truth = load_queries(root / kind / "queries.csv", rule=rule)
        got = cert.certify_positions(x, is_ee, "plain")
        _check(got, oracle.tuv_cert(is_ee, x), (rule, kind))
        assert got["hit"][truth].all() and list(got["hit"]) == list(truth), (kind, got["hit"], truth)
