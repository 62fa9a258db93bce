"""The allocation-free certificate (p2c_tuv_certify_static, scratch in a
static device array) against the entry points it is a variant of
(p2c_tuv_stream + p2c_tuv_wave + p2c_tuv_replay with the same parameters and
caller-owned scratch) and the oracle: identical hits and bit-identical TOI on
two calls in a row, the second smaller, so stale scratch from the first call
cannot leak into the second."""
import numpy as np
import pytest
import torch

import oracle
from p2cccd import _native as nat
from p2cccd import workloads
from p2cccd.tuv import TUV_RESULT_DTYPE, TuvCertifier, TuvConfig, results_view

pytestmark = pytest.mark.gpu


def _static_certify(cert: TuvCertifier, n: int, max_itr: int) -> np.ndarray:
    out = torch.zeros(n * TUV_RESULT_DTYPE.itemsize, dtype=torch.uint8, device=cert.device)
    nat.call("p2c_tuv_certify_static", cert.pairs.data_ptr(), n, cert.cfg.tolerance, max_itr, out.data_ptr(),
             nat.stream_handle())
    torch.cuda.synchronize(cert.device)
    return results_view(out)


def test_static_certificate_matches_stream_entry_points(gpu):
    assert nat.lib.p2c_tuv_static_max_pairs() == 1 << 18
    qs = workloads.make_query_set(20_000, 3_000, seed=77)
    # call 1: 20k VF pairs with a work cap of 40 boxes, so the wave pass and
    # the replay (heaps in the static array) both run; call 2: 3k EE pairs
    for is_ee, x, max_itr in ((False, qs.vf, 40), (True, qs.ee, 1_000_000)):
        cfg = TuvConfig(max_itr=max_itr, heap_cap=1 << 16, replay_slots=8, light_budget=1024, refill_at=8,
                        handover_at=8, blocks_per_cu=3, split_at=0, small_batch_split_at=0)
        cert = TuvCertifier(cfg, device=gpu)
        xd = torch.from_numpy(np.ascontiguousarray(x)).to(gpu)
        n = cert.setup_positions(xd, is_ee)
        print(f"static certificate: {'EE' if is_ee else 'VF'} n={n} max_itr={max_itr}", flush=True)
        got = _static_certify(cert, n, max_itr)
        cert.certify(n, "stream")
        torch.cuda.synchronize(gpu)
        ref = results_view(cert.result_tensor(n))
        assert np.array_equal(got["hit"], ref["hit"]) and np.array_equal(got["toi"], ref["toi"]), is_ee
        assert np.array_equal(got["flags"], ref["flags"]), is_ee
        o = oracle.tuv_cert(is_ee, x, max_itr=max_itr, heap_cap=1 << 16, threads=8)
        assert np.array_equal(got["hit"], o["hit"].astype(got["hit"].dtype)) and np.array_equal(got["toi"], o["toi"])
        if max_itr == 40:
            assert cert.replay_count() > 0


def test_static_certificate_refuses_oversized_batches(gpu):
    cert = TuvCertifier(device=gpu)
    cert.reserve(64)
    with pytest.raises(nat.NativeError, match="at most 262144 pairs"):
        nat.call("p2c_tuv_certify_static", cert.pairs.data_ptr(), (1 << 18) + 1, 1e-6, 1_000_000,
                 cert.results.data_ptr(), nat.stream_handle())
    nat.call("p2c_tuv_certify_static", cert.pairs.data_ptr(), 0, 1e-6, 1_000_000, cert.results.data_ptr(),
             nat.stream_handle())
