"""(t,u,v) inclusion certificate on the GPU: host-side driver over the C-ABI.

`TuvCertifier` owns the device workspaces (per-pair records, results, the
fallback-replay list and heaps) and runs one of four execution forms:
  plain       one lane per pair, one launch (the per-stage baseline)
  stream      chip-sized grid, one lane per pair; a wave refills its idle lanes
              together once `refill_at` of them are idle; pairs over
              `light_budget` boxes save their walk for the wave-cooperative pass
  rounds      budgeted one-lane-per-pair rounds that save unfinished walks and
              compact them, then the wave-cooperative pass
  persistent  chip-sized grid whose lanes refill one by one as pairs finish,
              then the wave-cooperative pass
All are followed by the fallback replay of pairs that exceeded max_itr, so
every pair ends with a final collide/separate decision.

Parameters mirror the reference's Tight-Inclusion call sites
(src/tools/tight_inclusion_full_query_benchmark.cpp:24-40): ms 0,
tolerance 1e-6, t_max 1, max_itr 1e6.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _native as nat

TUV_RESULT_DTYPE = np.dtype(
    [("toi", "<f8"), ("boxes", "<u8"), ("hit", "<u4"), ("flags", "<u4"), ("reserved", "<u8")]
)
FLAG_OVER_BUDGET, FLAG_REPLAYED, FLAG_CAPPED = 1, 2, 4
PAIR_BYTES = 288

_vp, _i64, _i32, _u32, _f64 = nat._vp, nat._i64, nat._i32, nat._u32, nat._f64
nat._sig("p2c_tuv_setup_positions", [_i32, _vp, _i64, _f64, _f64, _vp, _vp])
nat._sig("p2c_tuv_setup_indexed", [_i32, _vp, _vp, _vp, _i64, _f64, _f64, _vp, _vp])
nat._sig("p2c_tuv_setup_pairs", [_i32, _vp, _vp, _vp, _vp, _i64, _vp, _i64, _i64, _f64, _f64, _vp, _vp])
nat._sig("p2c_tuv_plain", [_vp, _i64, _vp, _f64, _i64, _vp, _vp, _vp, _u32, _vp])
nat._sig("p2c_tuv_persistent", [_vp, _i64, _vp, _f64, _i64, _i64, _vp, _vp, _vp, _vp, _u32, _i32, _vp])
nat._sig("p2c_tuv_stream", [_vp, _i64, _vp, _f64, _i64, _i64, _i32, _i32, _vp, _vp, _vp, _vp, _u32, _i32, _vp])
nat._sig("p2c_tuv_wave", [_vp, _vp, _vp, _vp, _u32, _f64, _i64, _vp, _vp, _i32, _vp])
nat._sig("p2c_tuv_wave_split", [_vp, _vp, _vp, _vp, _u32, _f64, _i64, _vp, _vp, _i64, _vp, _u32, _vp, _i32, _vp])
nat.lib.p2c_tuv_task_bytes.restype = ctypes.c_size_t
nat._sig("p2c_tuv_rounds", [_vp, _i64, _vp, _f64, _i64, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _i32, _vp])
nat.lib.p2c_tuv_state_bytes.restype = nat.ctypes.c_size_t
nat._sig("p2c_tuv_replay", [_vp, _vp, _vp, _u32, _f64, _i64, _vp, _i64, _i64, _vp, _vp])
nat.lib.p2c_tuv_pair_bytes.restype = nat.ctypes.c_size_t
nat._sig("p2c_tuv_certify_static", [_vp, _i64, _f64, _i64, _vp, _vp])
nat._sig("p2c_tuv_setup_pairs_lite", [_i32, _vp, _vp, _vp, _vp, _i64, _vp, _i64, _i64, _f64, _f64, _vp, _vp])
nat._sig("p2c_tuv_plain_lite", [_vp, _vp, _i64, _vp, _f64, _i64, _vp, _vp, _vp, _u32, _vp])
nat._sig("p2c_tuv_stream_lite", [_vp, _vp, _i64, _vp, _f64, _i64, _i64, _i32, _i32, _vp, _vp, _vp, _vp, _u32, _i32,
                                  _vp])
nat._sig("p2c_tuv_wave_split_lite", [_vp, _vp, _vp, _vp, _vp, _u32, _f64, _i64, _vp, _vp, _i64, _vp, _u32, _vp, _i32,
                                      _vp])
nat._sig("p2c_tuv_replay_lite", [_vp, _vp, _vp, _vp, _u32, _f64, _i64, _vp, _i64, _i64, _vp, _vp])
nat.lib.p2c_tuv_lite_bytes.argtypes = [nat._i64]
nat.lib.p2c_tuv_lite_bytes.restype = nat.ctypes.c_size_t
# p2c_tuv_vertices: the vertex arrays compact pair records are expanded from
VERTICES_DTYPE = np.dtype([("x0", "<u8"), ("x1", "<u8"), ("ms", "<f8")])
# p2c_tuv_scene: one scene of a batch's pair list (device topology addresses, first vertex row)
TUV_SCENE_DTYPE = np.dtype([("faces", "<u8"), ("edges", "<u8"), ("vertex_base", "<i8")])
nat._sig("p2c_tuv_setup_pairs_lite_batch", [nat._vp, nat._vp, nat._vp, nat._i32, nat._vp, nat._vp, nat._i64, nat._f64,
                                            nat._f64, nat._vp, nat._vp])
nat.lib.p2c_tuv_static_max_pairs.restype = nat.ctypes.c_size_t
nat.lib.p2c_tuv_pairs_bytes.argtypes = [nat._i64]
nat.lib.p2c_tuv_pairs_bytes.restype = nat.ctypes.c_size_t
assert nat.lib.p2c_tuv_pair_bytes() == PAIR_BYTES


@dataclass
class TuvConfig:
    tolerance: float = 1.0e-6
    ms: float = 0.0
    max_itr: int = 1_000_000
    heap_cap: int = 1 << 20      # canonical-replay heap entries per slot
    replay_slots: int = 8        # pairs replayed concurrently
    blocks_per_cu: int = 3       # persistent grid: resident blocks of 256 per CU
    light_budget: int = 1024     # boxes a pair may take in the stream/persistent pass before the wave pass
    refill_at: int = 8           # stream: idle lanes of a wave that trigger a refill
    handover_at: int = 8         # stream: busy lanes of a drained wave at or below which all go to the wave pass
    split_at: int = 0            # wave pass: boxes a wave takes on a pair before it hands walks to waiting waves
    small_batch_split_at: int = 4096  # split_at of a batch of at most one pair per grid lane when split_at is 0
                                      # (0: off); large batches keep their waves busy without splitting
    small_batch_budget: int = 0   # stream: light budget of a batch filling at most 1/8 of the lanes (0: off)
    split_tasks: int = 1 << 16   # capacity of the piece buffer
    round_budgets: tuple = (8, 32, 128, 512)  # walk steps per pair per round before the wave pass


def results_view(t: torch.Tensor) -> np.ndarray:
    return t.cpu().numpy().view(TUV_RESULT_DTYPE)


class TuvCertifier:
    """Certifies batches of VF or EE pairs.  Workspaces grow to the largest
    batch seen and are reused; nothing is allocated when a batch fits."""

    def __init__(self, config: TuvConfig | None = None, device: str | torch.device = "cuda"):
        self.cfg = config or TuvConfig()
        self.device = torch.device(device)
        self._cap = 0
        props = torch.cuda.get_device_properties(self.device)
        self.grid_blocks = int(props.multi_processor_count) * self.cfg.blocks_per_cu
        self.heaps = torch.empty(self.cfg.replay_slots * self.cfg.heap_cap * 32, dtype=torch.uint8,
                                 device=self.device)
        # u32 view: [0:2] next pair, [2] replay count, [3] heavy count, [4] heavy cursor, [5:13] round survivors
        # (stream form: [6:8] the EE range cursor),
        # [13] split pieces, [14] piece cursor, [15] split pairs
        self.counters = torch.zeros(8, dtype=torch.int64, device=self.device)
        self.budgets = np.ascontiguousarray(self.cfg.round_budgets, dtype=np.uint32)
        self.state_bytes = int(nat.lib.p2c_tuv_state_bytes())
        self.tasks = torch.empty(max(self.cfg.split_tasks, 1) * int(nat.lib.p2c_tuv_task_bytes()), dtype=torch.uint8,
                                 device=self.device)
        self._full = torch.empty(0, dtype=torch.uint8, device=self.device)
        self.verts = np.zeros(1, VERTICES_DTYPE)  # set by setup_pairs(lite=True)
        self.lite_active = False  # the last setup wrote compact records

    def reserve(self, n: int) -> None:
        self._reserve(n)

    @property
    def pairs(self) -> torch.Tensor:
        """The 288-byte records, allocated on first use: the compact form
        (64 bytes per pair) needs none."""
        need = int(nat.lib.p2c_tuv_pairs_bytes(self._cap))
        if self._full.numel() < need:
            self._full = torch.empty(need, dtype=torch.uint8, device=self.device)
        return self._full

    def _reserve(self, n: int) -> None:
        if n <= self._cap:
            return
        cap = (max(n, 1024) + 63) // 64 * 64  # pair records are stored in tiles of 64
        self.lite = torch.empty(int(nat.lib.p2c_tuv_lite_bytes(cap)), dtype=torch.uint8, device=self.device)
        self.results = torch.empty(cap * TUV_RESULT_DTYPE.itemsize, dtype=torch.uint8, device=self.device)
        self.replay_list = torch.empty(cap, dtype=torch.int32, device=self.device)
        self.heavy_list = torch.empty(cap, dtype=torch.int32, device=self.device)
        self.round_list = torch.empty(cap, dtype=torch.int32, device=self.device)
        self.states = torch.empty(cap * self.state_bytes, dtype=torch.uint8, device=self.device)
        self.split_list = torch.empty(cap, dtype=torch.int32, device=self.device)
        self._cap = cap

    # -- stages ------------------------------------------------------------
    def setup_positions(self, x: torch.Tensor, is_ee: bool, stream=None) -> int:
        assert x.dtype == torch.float64 and x.is_cuda and x.shape[1:] == (8, 3) and x.is_contiguous()
        n = x.shape[0]
        self._reserve(n)
        self.lite_active = False
        nat.call("p2c_tuv_setup_positions", int(is_ee), x.data_ptr(), n, self.cfg.ms, self.cfg.tolerance,
                 self.pairs.data_ptr(), nat.stream_handle(stream))
        return n

    def setup_indexed(self, x0: torch.Tensor, x1: torch.Tensor, idx4: torch.Tensor, is_ee: bool, stream=None) -> int:
        assert x0.dtype == torch.float64 and idx4.dtype == torch.int32 and idx4.shape[1] == 4
        n = idx4.shape[0]
        self._reserve(n)
        self.lite_active = False
        nat.call("p2c_tuv_setup_indexed", int(is_ee), x0.data_ptr(), x1.data_ptr(), idx4.data_ptr(), n,
                 self.cfg.ms, self.cfg.tolerance, self.pairs.data_ptr(), nat.stream_handle(stream))
        return n

    def certify(self, n: int, mode: str = "stream", stream=None, totals: torch.Tensor | None = None,
                events: tuple | None = None) -> None:
        """Certifies the pairs set up last (n, or min(n, totals[0] + totals[1])
        when device totals are given); results land in self.results.
        `events` (start, end[, mid]) are recorded around the certificate
        kernels (stream form: `mid` between the stream and the wave pass),
        before the replay."""
        s = nat.stream_handle(stream)
        if self.lite_active and mode not in ("stream", "plain"):
            raise ValueError("compact pair records are certified by the stream and plain forms only")
        c = self.counters
        c.zero_()
        if events is not None:
            events[0].record()
        rc_ptr = c.data_ptr() + 8  # replay count (uint32 in the low half of slot 1)
        tp = 0 if totals is None else totals.data_ptr()
        if mode == "plain":
            plain_args = (n, tp, self.cfg.tolerance, self.cfg.max_itr, self.results.data_ptr(),
                          self.replay_list.data_ptr(), rc_ptr, self._cap, s)
            if self.lite_active:
                nat.call("p2c_tuv_plain_lite", self.lite.data_ptr(), self.verts.ctypes.data, *plain_args)
            else:
                nat.call("p2c_tuv_plain", self.pairs.data_ptr(), *plain_args)
        elif mode == "rounds":
            nat.call("p2c_tuv_rounds", self.pairs.data_ptr(), n, tp, self.cfg.tolerance, self.cfg.max_itr,
                     self.budgets.ctypes.data, len(self.budgets), self.results.data_ptr(), c.data_ptr(),
                     self.states.data_ptr(), self.round_list.data_ptr(), self.heavy_list.data_ptr(),
                     self.replay_list.data_ptr(), self._cap, self.grid_blocks, s)
        elif mode == "stream":
            # optional: a batch that fills at most an eighth of the grid's lanes may use a smaller
            # light budget, moving its longest walks to the wave-shared pass at once
            budget = self.cfg.light_budget
            if self.cfg.small_batch_budget > 0 and n * 8 <= self.grid_blocks * 256:
                budget = min(budget, self.cfg.small_batch_budget)
            stream_args = (n, tp, self.cfg.tolerance, self.cfg.max_itr, budget, self.cfg.refill_at,
                           self.cfg.handover_at, self.results.data_ptr(), c.data_ptr(), self.states.data_ptr(),
                           self.heavy_list.data_ptr(), self._cap, self.grid_blocks, s)
            split_at = self.cfg.split_at
            if split_at == 0 and n <= self.grid_blocks * 256:
                split_at = self.cfg.small_batch_split_at
            wave_args = (c.data_ptr(), self.states.data_ptr(), self.heavy_list.data_ptr(), self._cap,
                         self.cfg.tolerance, self.cfg.max_itr, self.results.data_ptr(), self.replay_list.data_ptr(),
                         split_at, self.tasks.data_ptr(), self.cfg.split_tasks, self.split_list.data_ptr(),
                         self.grid_blocks, s)
            if self.lite_active:
                nat.call("p2c_tuv_stream_lite", self.lite.data_ptr(), self.verts.ctypes.data, *stream_args)
            else:
                nat.call("p2c_tuv_stream", self.pairs.data_ptr(), *stream_args)
            if events is not None and len(events) > 2:
                events[2].record()
            if self.lite_active:
                nat.call("p2c_tuv_wave_split_lite", self.lite.data_ptr(), self.verts.ctypes.data, *wave_args)
            else:
                nat.call("p2c_tuv_wave_split", self.pairs.data_ptr(), *wave_args)
        elif mode == "persistent":
            nat.call("p2c_tuv_persistent", self.pairs.data_ptr(), n, tp, self.cfg.tolerance, self.cfg.max_itr,
                     self.cfg.light_budget, self.results.data_ptr(), c.data_ptr(), self.heavy_list.data_ptr(),
                     self.replay_list.data_ptr(), self._cap, self.grid_blocks, s)
        else:
            raise ValueError(f"unknown certificate mode {mode!r}")
        if events is not None:
            events[1].record()
        replay_args = (self.replay_list.data_ptr(), rc_ptr, self._cap, self.cfg.tolerance, self.cfg.max_itr,
                       self.heaps.data_ptr(), self.cfg.heap_cap, self.cfg.replay_slots, self.results.data_ptr(), s)
        if self.lite_active:
            nat.call("p2c_tuv_replay_lite", self.lite.data_ptr(), self.verts.ctypes.data, *replay_args)
        else:
            nat.call("p2c_tuv_replay", self.pairs.data_ptr(), *replay_args)

    def setup_pairs(self, x0: torch.Tensor, x1: torch.Tensor, topo: torch.Tensor, pairs: torch.Tensor, is_ee: bool,
                    n: int, totals: torch.Tensor | None = None, out_offset: int = 0, stream=None,
                    lite: bool = False) -> None:
        """Pair records straight from broad-phase pair lists (see
        p2c_tuv_setup_pairs).  lite: 64-byte records expanded from x0 and x1
        when loaded (p2c_tuv_setup_pairs_lite); every pair of the batch must
        then index the same x0 and x1, and the batch is certified by the
        stream form."""
        args = (int(is_ee), x0.data_ptr(), x1.data_ptr(), topo.data_ptr(), pairs.data_ptr(), n,
                0 if totals is None else totals.data_ptr(), out_offset, self._cap, self.cfg.ms, self.cfg.tolerance)
        if lite:
            self.verts["x0"], self.verts["x1"], self.verts["ms"] = x0.data_ptr(), x1.data_ptr(), self.cfg.ms
            nat.call("p2c_tuv_setup_pairs_lite", *args, self.lite.data_ptr(), nat.stream_handle(stream))
        else:
            nat.call("p2c_tuv_setup_pairs", *args, self.pairs.data_ptr(), nat.stream_handle(stream))
        self.lite_active = lite

    def setup_batch_lite(self, x0: torch.Tensor, x1: torch.Tensor, scenes: torch.Tensor, n_scenes: int,
                         pairs: torch.Tensor, offsets: torch.Tensor, n: int, stream=None) -> None:
        """Compact records of a batch of scenes' pair list (the pairs and
        offsets of BatchBroadPhase; p2c_tuv_setup_pairs_lite_batch): x0 and
        x1 hold every scene's vertices, `scenes` is the device table of
        TUV_SCENE_DTYPE rows (faces, edges, first vertex row)."""
        self._reserve(n)
        self.verts["x0"], self.verts["x1"], self.verts["ms"] = x0.data_ptr(), x1.data_ptr(), self.cfg.ms
        nat.call("p2c_tuv_setup_pairs_lite_batch", x0.data_ptr(), x1.data_ptr(), scenes.data_ptr(), n_scenes,
                 pairs.data_ptr(), offsets.data_ptr(), n, self.cfg.ms, self.cfg.tolerance, self.lite.data_ptr(),
                 nat.stream_handle(stream))
        self.lite_active = True

    def replay_count(self) -> int:
        return int(self.counters[1].item() & 0xFFFFFFFF)

    def split_counts(self) -> tuple:
        """(pieces written, pairs split) by the last stream-form wave pass."""
        c = self.counters.cpu().numpy().view(np.uint32)
        return int(c[13]) + int(c[8]), int(c[15])

    def heavy_count(self) -> int:
        return int((self.counters[1].item() >> 32) & 0xFFFFFFFF)

    def result_tensor(self, n: int) -> torch.Tensor:
        return self.results[: n * TUV_RESULT_DTYPE.itemsize]

    # -- convenience -------------------------------------------------------
    def certify_positions(self, x, is_ee: bool, mode: str = "stream") -> np.ndarray:
        if isinstance(x, np.ndarray):
            x = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(self.device)
        n = self.setup_positions(x, is_ee)
        self.certify(n, mode)
        torch.cuda.synchronize(self.device)
        return results_view(self.result_tensor(n))
