"""Serving: per-timestep CCD requests against many resident scenes.

Each resident scene keeps its topology and its broad-phase state (boxes,
object envelopes and pair matrix, both LBVHs) in HBM.  A request carries one
timestep's positions (x0, x1) for one scene.  The server runs a
continuous-batching loop over two lanes (a HIP stream, a batch broad-phase
workspace, batch position and pair buffers and a certificate workspace
each):

  * admission: a queued request joins the next batch when its scene is
    resident or can be made resident within the HBM budget (least recently
    used idle scenes are evicted), its scene has no request in flight, and
    the batch has request and pair slots left;
  * a batch uploads every request's positions in one copy, builds and counts
    every scene's trees with one launch per stage for the whole batch
    (BatchBroadPhase), reads the pair offsets with one wait on its lane,
    emits one pair list (scene after scene), sets up compact certificate
    records for all of it in one launch and certifies them in one shared
    launch (stream kernel + wave pass + replay), so requests of mixed size
    share every launch;
  * while one lane's batch is on the GPU, the host admits new arrivals into
    the other lane's batch, whose broad phase overlaps the first batch's
    certificate on the device.

The HBM budget covers the resident scenes; each lane's working buffers are
sized by the largest batch it has run.

Per-pair answers do not depend on what else shares a batch: candidate order
is fixed per scene and the certificate's result per pair is order-free, so a
request's result is bit-identical alone and in the mix (tests check it).

The workload of the serving benchmark: resident scenes (256 by default),
each a pair of jittered icospheres or grid meshes with log-uniform triangle
counts (1k-500k); Poisson arrivals; each request one step of random rigid
motion plus per-vertex N(0, 1e-3) noise.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import scenes
from .broadphase import BatchBroadPhase, ResidentMesh
from .tuv import TUV_RESULT_DTYPE, TUV_SCENE_DTYPE, TuvCertifier, TuvConfig


@dataclass
class Request:
    rid: int
    scene_id: int
    x0: np.ndarray
    x1: np.ndarray
    arrival: float = 0.0


@dataclass
class Completed:
    """A served timestep: the candidate counts, every certified contact (its
    pair and TOI, in candidate order) and the earliest TOI.  With the
    server's full_results, also every candidate pair and its hit flag and TOI."""
    rid: int
    scene_id: int
    arrival: float
    done: float
    n_vf: int
    n_ee: int
    vf_hits: np.ndarray  # int32 (k, 2): (vertex, face) pairs with a contact
    vf_toi: np.ndarray   # float64 (k,)
    ee_hits: np.ndarray  # int32 (m, 2): (edge, edge) pairs with a contact
    ee_toi: np.ndarray   # float64 (m,)
    full: dict | None = None  # vf, ee (all candidates), hit, toi (per candidate, VF then EE)

    @property
    def latency_ms(self) -> float:
        return (self.done - self.arrival) * 1e3

    @property
    def pairs(self) -> int:
        return self.n_vf + self.n_ee

    @property
    def earliest_toi(self) -> float:
        t = np.concatenate([self.vf_toi, self.ee_toi])
        return float(t.min()) if len(t) else float("inf")


# ---- admission (host logic only) ------------------------------------------------
class Residency:
    """Which scenes hold HBM, under a byte budget, with LRU eviction of scenes
    that have no request in flight."""

    def __init__(self, budget: int):
        self.budget = budget
        self.bytes: dict[int, int] = {}
        self.lru: collections.OrderedDict[int, None] = collections.OrderedDict()
        self.busy: set[int] = set()
        self.evictions = 0

    @property
    def used(self) -> int:
        return sum(self.bytes.values())

    def touch(self, sid: int) -> None:
        self.lru.pop(sid, None)
        self.lru[sid] = None

    def plan(self, sid: int, need: int) -> list[int] | None:
        """Scenes to evict so that `sid` (needing `need` bytes if not resident)
        fits; None if it cannot fit now."""
        if sid in self.bytes:
            return []
        if need > self.budget:
            raise RuntimeError(f"scene {sid} needs {need} bytes, more than the HBM budget {self.budget}")
        free = self.budget - self.used
        out = []
        for victim in self.lru:
            if free >= need:
                break
            if victim in self.busy or victim == sid:
                continue
            out.append(victim)
            free += self.bytes[victim]
        return out if free >= need else None

    def add(self, sid: int, nbytes: int) -> None:
        self.bytes[sid] = nbytes
        self.touch(sid)

    def remove(self, sid: int) -> None:
        self.bytes.pop(sid, None)
        self.lru.pop(sid, None)
        self.evictions += 1


# ---- device state -------------------------------------------------------------
CACHE_SLOTS = 16  # hit-cache partners per query of a batch of sparse scenes (most of their queries have fewer)
CACHE_BYTES = 1 << 30  # the batch hit cache grows with the batch's pairs per query up to this size


def cache_slots(queries: int, pairs_estimate: int, base: int = CACHE_SLOTS) -> int:
    """Hit-cache partners per query for a batch: doubled from `base` while
    under 4x the batch's pairs per query (the count pass then keeps nearly
    every query's partners and the fill re-traverses few), at most 256 and at
    most CACHE_BYTES of cache."""
    slots = base
    while slots < 256 and slots < 4 * pairs_estimate / max(queries, 1) and 4 * queries * slots * 2 <= CACHE_BYTES:
        slots *= 2
    return slots


class ResidentScene:
    """A scene's topology and broad-phase state on the device."""

    def __init__(self, scene: scenes.Scene, device, host: dict | None = None):
        self.scene = scene
        self.mesh = ResidentMesh(scene.topo, scene.n_vertices, device, host)

    @staticmethod
    def estimate_bytes(scene: scenes.Scene) -> int:
        return ResidentMesh.footprint(scene.topo, scene.n_vertices)

    def nbytes(self) -> int:
        return self.mesh.nbytes()


@dataclass
class _Batch:
    reqs: list
    offsets: list = field(default_factory=list)  # first pair of each request
    counts: list = field(default_factory=list)   # (n_vf, n_ee) of each request
    total: int = 0
    event: torch.cuda.Event | None = None


def _grow(t: torch.Tensor, n: int) -> torch.Tensor:
    return t if t.numel() >= n else torch.empty(int(n * 1.25) + 1024, dtype=t.dtype, device=t.device)


class Lane:
    """A stream and the working buffers of the batch it runs."""

    def __init__(self, device, cfg: TuvConfig | None, cache_slots: int):
        self.device = device
        self.stream = torch.cuda.Stream(device)
        # lanes' kernels overlap on the device, and a splitting wave pass keeps its waves resident while they
        # wait for pieces: by default a lane's batches are not split
        self.cert = TuvCertifier(cfg if cfg is not None else TuvConfig(small_batch_split_at=0), device)
        self.bp = BatchBroadPhase(device, cache_slots)
        self.batch: _Batch | None = None
        self.stage = torch.empty(0, dtype=torch.float64).pin_memory()  # pinned staging of the batch's positions
        self.tab_host = torch.empty(0, dtype=torch.uint8).pin_memory()  # pinned staging of the setup's scene table
        self.xb = torch.empty(0, dtype=torch.float64, device=device)  # the batch's x0 rows, then its x1 rows
        self.tab = torch.empty(0, dtype=torch.uint8, device=device)
        self.pairs = torch.empty((0, 2), dtype=torch.int32, device=device)

    def staging(self, n: int) -> torch.Tensor:
        if self.stage.numel() < n:
            self.stage = torch.empty(int(n * 1.25) + 1024, dtype=torch.float64).pin_memory()
        return self.stage

    def table_staging(self, n: int) -> torch.Tensor:
        if self.tab_host.numel() < n:
            self.tab_host = torch.empty(int(n * 1.25) + 1024, dtype=torch.uint8).pin_memory()
        return self.tab_host

    def reserve(self, n_pos: int, n_tab: int, n_pairs: int) -> None:
        """Device buffers for a batch (call on the lane's stream)."""
        self.xb = _grow(self.xb, n_pos)
        self.tab = _grow(self.tab, n_tab)
        if self.pairs.shape[0] < n_pairs:
            self.pairs = torch.empty((int(n_pairs * 1.25) + 1024, 2), dtype=torch.int32, device=self.device)


class Server:
    def __init__(self, scene_table: dict[int, scenes.Scene], hbm_budget: int, device="cuda",
                 max_batch_requests: int = 64, max_batch_pairs: int = 1 << 23, cfg: TuvConfig | None = None,
                 lanes: int = 2, full_results: bool = False, cache_slots: int = CACHE_SLOTS):
        self.full_results = full_results
        self.table = scene_table
        self.device = torch.device(device)
        self.res = Residency(hbm_budget)
        self.resident: dict[int, ResidentScene] = {}
        self.last_pairs: dict[int, int] = {}
        self.max_batch_requests = max_batch_requests
        self.max_batch_pairs = max_batch_pairs
        self.lanes = [Lane(self.device, cfg, cache_slots) for _ in range(lanes)]
        self.batches = 0
        self.loads = 0
        self._host_topo: dict[int, dict] = {}  # pinned topology of every scene loaded so far

    # -- residency ---------------------------------------------------------
    def _make_resident(self, sid: int, stream) -> bool:
        """Resident or loaded now (topology from a pinned host copy, uploaded
        on `stream`, the stream of the lane whose batch needs it)."""
        if sid in self.resident:
            self.res.touch(sid)
            return True
        need = ResidentScene.estimate_bytes(self.table[sid])
        plan = self.res.plan(sid, need)
        if plan is None:
            return False
        for victim in plan:
            del self.resident[victim]
            self.res.remove(victim)
        host = self._host_topo.get(sid)
        if host is None:
            host = self._host_topo[sid] = ResidentMesh.pinned_topology(self.table[sid].topo)
        with torch.cuda.stream(stream):
            self.resident[sid] = ResidentScene(self.table[sid], self.device, host)
        self.res.add(sid, need)
        self.loads += 1
        return True

    # -- batches -----------------------------------------------------------
    def _admit(self, queue: collections.deque, stream) -> list:
        batch, pairs, skipped = [], 0, collections.deque()
        busy_now = set()
        while queue and len(batch) < self.max_batch_requests:
            r = queue.popleft()
            # pairs of the scene's last step, or 16 per triangle before its first
            est = self.last_pairs.get(r.scene_id, 16 * len(self.table[r.scene_id].topo.faces))
            if r.scene_id in self.res.busy or r.scene_id in busy_now or (batch and pairs + est > self.max_batch_pairs):
                skipped.append(r)
                continue
            if not self._make_resident(r.scene_id, stream):
                skipped.append(r)
                continue
            batch.append(r)
            busy_now.add(r.scene_id)
            self.res.busy.add(r.scene_id)
            pairs += est
        queue.extendleft(reversed(skipped))
        return batch

    def _launch(self, lane: Lane, reqs: list) -> None:
        b = _Batch(reqs=reqs)
        S = len(reqs)
        nvs = [self.table[r.scene_id].n_vertices for r in reqs]
        base = np.concatenate([[0], np.cumsum(nvs)]).astype(np.int64)
        V = int(base[-1])
        # positions through the lane's pinned staging: every x0 row, then every x1 row, one copy
        stage = lane.staging(6 * V)
        host = stage.numpy()
        for k, r in enumerate(reqs):
            host[3 * base[k]: 3 * base[k + 1]] = r.x0.reshape(-1)
            host[3 * (V + base[k]): 3 * (V + base[k + 1])] = r.x1.reshape(-1)
        tab = np.zeros(S, TUV_SCENE_DTYPE)
        s = lane.stream
        with torch.cuda.stream(s):
            lane.reserve(6 * V, tab.nbytes, 0)
            lane.xb[: 6 * V].copy_(stage[: 6 * V], non_blocking=True)
            x0, x1 = lane.xb[: 3 * V], lane.xb[3 * V: 6 * V]
            items = []
            for k, r in enumerate(reqs):
                rs = self.resident[r.scene_id]
                items.append((rs.mesh, x0.data_ptr() + 24 * int(base[k]), x1.data_ptr() + 24 * int(base[k]),
                              rs.scene.thickness))
                tab[k] = (rs.mesh.faces.data_ptr(), rs.mesh.edges.data_ptr(), base[k])
            queries = sum(self.table[r.scene_id].n_vertices + len(self.table[r.scene_id].topo.edges) for r in reqs)
            est = sum(self.last_pairs.get(r.scene_id, 0) for r in reqs)
            offs_dev = lane.bp.count(items, stream=s, cache_slots=cache_slots(queries, est, lane.bp.cache_slots))
            offs = offs_dev.cpu().numpy().astype(np.int64)  # the counts size this batch; waits on this lane only
        if offs[2 * S + 1]:
            raise RuntimeError(f"{offs[2 * S + 1]} BVH traversals exceeded their step bound")
        for k, r in enumerate(reqs):
            n_vf, n_ee = int(offs[2 * k + 1] - offs[2 * k]), int(offs[2 * k + 2] - offs[2 * k + 1])
            b.offsets.append(int(offs[2 * k]))
            b.counts.append((n_vf, n_ee))
            self.last_pairs[r.scene_id] = n_vf + n_ee
        b.total = total = int(offs[2 * S])
        cert = lane.cert
        th = lane.table_staging(tab.nbytes)
        th.numpy()[: tab.nbytes] = tab.view(np.uint8)
        with torch.cuda.stream(s):
            lane.reserve(0, tab.nbytes, total)
            cert.reserve(max(total, 1))
            lane.bp.emit(lane.pairs, stream=s)
            if total:
                lane.tab[: tab.nbytes].copy_(th[: tab.nbytes], non_blocking=True)
                cert.setup_batch_lite(x0, x1, lane.tab, S, lane.pairs, offs_dev, total, stream=s)
                cert.certify(total, "stream", stream=s)
            b.event = torch.cuda.Event()
            b.event.record(s)
        lane.batch = b
        self.batches += 1

    def _complete(self, lane: Lane, out: list) -> None:
        """Reads a finished batch back: only the contacts (pair, TOI) leave
        the device, gathered there from the per-pair results."""
        b = lane.batch
        b.event.synchronize()
        with torch.cuda.stream(lane.stream):
            res = lane.cert.result_tensor(max(b.total, 1)).view(torch.float64).view(-1, 4)[: b.total]
            hit = (res[:, 2].view(torch.int64) & 0xFFFFFFFF) != 0
            idx = torch.nonzero(hit).squeeze(1)  # ascending: requests' VF then EE ranges in order
            ends = [e for o, (n_vf, n_ee) in zip(b.offsets, b.counts) for e in (o, o + n_vf, o + n_vf + n_ee)]
            cut = torch.searchsorted(idx, torch.tensor(ends, dtype=torch.int64, device=idx.device)).cpu().tolist()
            hit_pairs = lane.pairs[idx].to("cpu", non_blocking=True)
            hit_toi = res[idx, 0].to("cpu", non_blocking=True)
            full = None
            if self.full_results:
                full = [(lane.pairs[o:o + n_vf].cpu().numpy(), lane.pairs[o + n_vf:o + n_vf + n_ee].cpu().numpy(),
                         res[o:o + n_vf + n_ee].cpu().numpy()) for o, (n_vf, n_ee) in zip(b.offsets, b.counts)]
            lane.stream.synchronize()
        now = time.perf_counter()
        hp, ht = hit_pairs.numpy(), hit_toi.numpy()
        for k, (r, (n_vf, n_ee)) in enumerate(zip(b.reqs, b.counts)):
            lo, mid, hi = cut[3 * k: 3 * k + 3]
            extra = None
            if full is not None:
                vf, ee, rr = full[k]
                view = rr.view(TUV_RESULT_DTYPE).reshape(-1)
                extra = {"vf": vf, "ee": ee, "hit": view["hit"].copy(), "toi": view["toi"].copy()}
            out.append(Completed(r.rid, r.scene_id, r.arrival, now, n_vf, n_ee, hp[lo:mid], ht[lo:mid], hp[mid:hi],
                                 ht[mid:hi], extra))
            self.res.busy.discard(r.scene_id)
        lane.batch = None

    def run(self, requests: list, realtime: bool = True) -> list:
        """Serves `requests` (arrival offsets in seconds from the start when
        realtime, otherwise all queued at once); returns completions with
        arrival/done on the same clock (perf_counter)."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue: collections.deque = collections.deque()
        out: list = []
        t0 = time.perf_counter()
        launched: list = []  # busy lanes, oldest first
        while pending or queue or launched:
            now = time.perf_counter() - t0
            while pending and (not realtime or pending[0].arrival <= now):
                r = pending.popleft()
                if not realtime:
                    r.arrival = now
                queue.append(r)
            for lane in list(launched):  # retire finished batches
                if lane.batch.event.query():
                    self._complete(lane, out)
                    launched.remove(lane)
            free = [lane for lane in self.lanes if lane.batch is None]
            if free and queue:
                reqs = self._admit(queue, free[0].stream)
                if reqs:
                    self._launch(free[0], reqs)
                    launched.append(free[0])
                    continue
            if launched and (queue or not pending):
                # nothing admissible right now: wait for the oldest batch
                self._complete(launched[0], out)
                launched.pop(0)
            elif pending:
                time.sleep(max(0.0, min(pending[0].arrival - (time.perf_counter() - t0), 0.001)))
            elif queue:
                raise RuntimeError("a queued request can never be admitted (HBM budget too small?)")
        for c in out:
            c.arrival += t0
        return out


# ---- the serving workload -------------------------------------------------------
def _grid_mesh(n: int, size: float):
    """n x n vertex grid in the z = 0 plane, 2 (n-1)^2 triangles."""
    u = np.linspace(-size / 2, size / 2, n)
    X, Y = np.meshgrid(u, u, indexing="ij")
    v = np.stack([X.ravel(), Y.ravel(), np.zeros(n * n)], 1)
    i = np.arange(n - 1)
    a = (i[:, None] * n + i[None, :]).ravel()
    f = np.concatenate([np.stack([a, a + n, a + 1], 1), np.stack([a + 1, a + n, a + n + 1], 1)])
    return v, f.astype(np.int32)


def make_scene(sid: int, triangles: int, seed: int) -> scenes.Scene:
    """Two jittered meshes of about `triangles` triangles in total, 0.02
    apart: icospheres (for about half of the seeds, when a subdivision level
    matches the size to within 2x) or square grids."""
    rng = np.random.default_rng(seed)
    per = max(triangles // 2, 20)
    level = int(round(np.log(per / 20.0) / np.log(4.0)))
    if rng.random() < 0.5 and level >= 1 and 0.5 <= 20 * 4 ** level / per <= 2.0:
        v, f = scenes.icosphere(level)
        v2 = v.copy() + np.array([2.02, 0.0, 0.0])
        kind = "spheres"
    else:
        n = max(int(np.ceil(np.sqrt(per / 2.0))) + 1, 3)
        v, f = _grid_mesh(n, 2.0)
        v2 = v.copy() + np.array([0.0, 0.0, 0.02])
        kind = "grids"
    v = v + rng.normal(0, 1e-3, v.shape)
    v2 = v2 + rng.normal(0, 1e-3, v2.shape)
    verts = np.concatenate([v, v2])
    faces = np.concatenate([f, f + len(v)]).astype(np.int32)
    topo = scenes.prepare_topology(faces, len(verts))
    return scenes.Scene(name=f"serve{sid}_{kind}_{len(faces)}", x0=verts, x1=verts.copy(), topo=topo,
                        thickness=0.0, info={"kind": kind, "triangles": int(len(faces))})


def step_request(rid: int, sid: int, scene: scenes.Scene, seed: int, arrival: float = 0.0) -> Request:
    """One timestep: the second mesh moves rigidly toward the first (random
    rotation up to 1 degree, translation 0.015 against a 0.02 gap), every
    vertex adds N(0, 1e-3)."""
    rng = np.random.default_rng(seed)
    x0 = scene.x0.copy()
    x1 = x0.copy()
    half = len(x0) // 2
    ang = np.deg2rad(rng.uniform(0, 1))
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    R = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * (K @ K)
    c = x0[half:].mean(0)
    direction = x0[:half].mean(0) - c
    direction /= np.linalg.norm(direction) + 1e-300
    x1[half:] = (x0[half:] - c) @ R.T + c + 0.015 * direction
    x1 += rng.normal(0, 1e-3, x1.shape)
    return Request(rid, sid, np.ascontiguousarray(x0), np.ascontiguousarray(x1), arrival)


def poisson_workload(n_scenes: int, n_requests: int, rate_hz: float, min_tri: int = 1000, max_tri: int = 500_000,
                     seed: int = 5):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(min_tri), np.log(max_tri), n_scenes)).astype(int)
    table = {i: make_scene(i, int(sizes[i]), seed + 7 * i) for i in range(n_scenes)}
    arrivals = np.cumsum(rng.exponential(1.0 / rate_hz, n_requests))
    sids = rng.integers(0, n_scenes, n_requests)
    reqs = [step_request(k, int(s), table[int(s)], seed + 1000 + k, float(a)) for k, (s, a) in
            enumerate(zip(sids, arrivals))]
    return table, reqs


def poisson_requests(table: dict, n_requests: int, rate_hz: float, seed: int) -> list:
    """More requests over an existing scene table: Poisson arrivals at
    rate_hz, scenes drawn uniformly."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate_hz, n_requests))
    ids = sorted(table)
    sids = rng.integers(0, len(ids), n_requests)
    return [step_request(k, ids[int(s)], table[ids[int(s)]], seed + 1000 + k, float(a)) for k, (s, a) in
            enumerate(zip(sids, arrivals))]


def footprint(table: dict) -> int:
    """Device bytes of every scene resident at once."""
    return sum(ResidentScene.estimate_bytes(s) for s in table.values())


def serve_timed(srv: Server, reqs: list) -> dict:
    """Serves `reqs` in real time on a warmed server; the report plus the
    batches, scene loads and evictions of this run."""
    b, loads, ev = srv.batches, srv.loads, srv.res.evictions
    t = time.perf_counter()
    done = srv.run(reqs, realtime=True)
    wall = time.perf_counter() - t
    rep = report(done, wall)
    rep.update(batches=srv.batches - b, scene_loads_during_run=srv.loads - loads, evictions=srv.res.evictions - ev,
               wall_s=wall)
    return rep


def report(done: list, wall: float) -> dict:
    lat = np.array([c.latency_ms for c in done])
    pairs = int(sum(c.pairs for c in done))
    return {"requests": len(done), "pairs": pairs, "pairs_per_s": pairs / wall, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean())}


def main() -> int:
    import argparse
    import json

    ap = argparse.ArgumentParser(description="serving benchmark: Poisson requests over resident scenes")
    ap.add_argument("--scenes", type=int, default=256)
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--rate", type=float, default=400.0, help="arrivals per second")
    ap.add_argument("--max-triangles", type=int, default=500_000)
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--budget-frac", type=float, default=0.0,
                    help="HBM budget as a fraction of every scene's resident footprint (overrides --budget-gb)")
    args = ap.parse_args()
    t = time.perf_counter()
    table, reqs = poisson_workload(args.scenes, args.requests, args.rate, max_tri=args.max_triangles)
    fp = footprint(table)
    budget = int(args.budget_frac * fp) if args.budget_frac > 0 else int(args.budget_gb * (1 << 30))
    print(f"[serve] built {len(table)} scenes in {time.perf_counter() - t:.1f} s; footprint {fp / 2**30:.2f} GiB, "
          f"budget {budget / 2**30:.2f} GiB", flush=True)
    srv = Server(table, budget)
    # make every scene resident (as far as the budget allows) and size the lanes with one step each
    warm = [step_request(-1 - i, sid, table[sid], 99 + i) for i, sid in enumerate(sorted(table))]
    srv.run(warm, realtime=False)
    print(f"[serve] {len(srv.resident)} scenes resident ({srv.res.used / 2**30:.2f} GiB)", flush=True)
    rep = serve_timed(srv, reqs)
    print(f"[serve] served {rep['requests']} requests in {rep['wall_s']:.1f} s", flush=True)
    rep.update(scenes=args.scenes, rate_hz=args.rate, budget_gb=budget / 2**30, footprint_gb=fp / 2**30)
    print(json.dumps(rep), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
