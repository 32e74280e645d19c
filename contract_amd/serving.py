"""Serving loop: contraction requests against propagators resident in HBM.

Requests (meson two-point grids, baryon two-points, diquarks) arrive over
time and reference propagators by key.  The engine keeps propagators in an
LRU pool with a byte cap, admits a waiting request when a launch slot is free
and the pool can hold everything the request needs (its missing inputs and
its outputs) without evicting anything in use, and batches continuously:
every scheduler tick, the meson requests admitted together go through one
batched persistent launch (contract_meson_grid_project), the baryon requests
through one batched fused launch (contract_baryon_twopoint_project_batch),
diquark requests get a launch each, and launches are spread over several
streams so that work admitted while earlier work is still on the GPU
overlaps with it.
A request completes when the event recorded after its last kernel has
passed; its latency is completion minus arrival.

The results do not depend on what else is in flight: each request's kernels
see the same inputs and the same work decomposition alone and in the mix (the
batched launches compute every request from its own work items and reduce
them in a fixed order), which the tests check.

Nothing here runs on the CPU in place of the GPU: every contraction is a
libcontract.so launch.
"""

from __future__ import annotations

import argparse
import dataclasses
import json
import sys
import time
import zlib
from collections import OrderedDict, deque
from typing import Callable, Dict, Hashable, List, Optional, Sequence, Tuple

import numpy as np

from . import pycontract as pc

KINDS = ("meson", "baryon", "diquark")
MOM27 = [(x, y, z) for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)]
PROTON = [(0, 1.0), (5, 1.0)]  # IK_JL_NM + IM_JL_NK on (u, d, u)
CG5 = 5
SITE_BYTES = {"complex128": 2304, "complex64": 1152}
GRID_TUNING = {"complex128": (1024, 4, 2), "complex64": (1024, 8, 2)}


@dataclasses.dataclass(frozen=True)
class PropKey:
    """A propagator of the catalogue: synthetic, reproducible from its fields
    (iid N(0,1) entries times exp(-alpha |t - t0|))."""
    dims: Tuple[int, int, int, int]
    index: int
    t0: int = 0
    alpha: float = 0.25
    dtype: str = "complex128"

    @property
    def nbytes(self) -> int:
        return int(np.prod(self.dims)) * SITE_BYTES[self.dtype]

    @property
    def seed(self) -> int:
        return zlib.crc32(repr((self.dims, self.index, self.t0, self.alpha, self.dtype)).encode()) & 0x7FFFFFFF


@dataclasses.dataclass(frozen=True)
class OutKey:
    """A propagator-shaped result held in the pool (a diquark)."""
    rid: int
    dims: Tuple[int, int, int, int]
    dtype: str = "complex128"

    @property
    def nbytes(self) -> int:
        return int(np.prod(self.dims)) * SITE_BYTES[self.dtype]


@dataclasses.dataclass
class Request:
    """kind 'meson': props (P_i, P_j), gammas = [(gamma_ij, gamma_kl)], result
    [len(gammas)][len(momenta)][Lt] = projected meson_two_point.
    kind 'baryon': props (u, d, u'), gammas = insertions Gamma_mn, result
    [len(gammas)][len(momenta)][Lt] = projected proton two-point (Cg5).
    kind 'diquark': props (P_i, P_j), gammas = [(gamma_ij, gamma_kl)], result
    = baryon_diquark output kept in the pool under OutKey(rid)."""
    rid: int
    kind: str
    dims: Tuple[int, int, int, int]
    props: Tuple[PropKey, ...]
    gammas: list
    momenta: list
    arrival: float = 0.0
    admitted: Optional[float] = None
    done: Optional[float] = None
    result: object = None

    @property
    def latency(self) -> Optional[float]:
        return None if self.done is None else self.done - self.arrival

    def out_key(self) -> Optional[OutKey]:
        return OutKey(self.rid, self.dims, self.props[0].dtype) if self.kind == "diquark" else None


class PoolFull(RuntimeError):
    pass


class PropagatorPool:
    """LRU pool of device propagators under a byte cap.

    acquire() pins an entry (materialising it if absent, after evicting
    least-recently-used unpinned entries); release() unpins it.  Pinned
    entries are never evicted, so a request's inputs stay put until its
    kernels have finished."""

    def __init__(self, cap_bytes: int, materialise: Callable[[Hashable], object]):
        self.cap = int(cap_bytes)
        self.materialise = materialise
        self.entries: "OrderedDict[Hashable, list]" = OrderedDict()  # key -> [value, nbytes, pins]
        self.used = 0
        self.hits = self.misses = self.evictions = 0
        self.peak = 0

    def __contains__(self, key) -> bool:
        return key in self.entries

    def pinned_bytes(self) -> int:
        return sum(e[1] for e in self.entries.values() if e[2] > 0)

    def fits(self, keys: Sequence[Hashable], extra: Sequence[Hashable] = ()) -> bool:
        """Could `keys` (inputs, kept if resident) plus the new entries `extra`
        be pinned together without evicting a pinned entry?"""
        need = sum(k.nbytes for k in set(keys) if k not in self.entries) + sum(k.nbytes for k in extra)
        kept = sum(self.entries[k][1] for k in set(keys) if k in self.entries and self.entries[k][2] == 0)
        return self.pinned_bytes() + kept + need <= self.cap

    def _make_room(self, nbytes: int):
        """Evict least-recently-used unpinned entries until `nbytes` fit.
        Returns the value of an evicted entry of exactly `nbytes` (its memory
        can be reused, which spares the allocator a free and a fresh
        allocation) or None."""
        if nbytes > self.cap:
            raise PoolFull(f"entry of {nbytes} bytes exceeds the pool cap of {self.cap}")
        recycled = None
        for key in list(self.entries):
            if self.used + nbytes <= self.cap:
                break
            value, size, pins = self.entries[key]
            if pins == 0:
                del self.entries[key]
                self.used -= size
                self.evictions += 1
                if recycled is None and size == nbytes and value is not None:
                    recycled = value
        if self.used + nbytes > self.cap:
            raise PoolFull(f"no room for {nbytes} bytes: {self.used} used, {self.pinned_bytes()} pinned")
        return recycled

    def acquire(self, key):
        e = self.entries.get(key)
        if e is not None:
            self.hits += 1
            self.entries.move_to_end(key)
            e[2] += 1
            return e[0]
        self.misses += 1
        recycled = self._make_room(key.nbytes)
        value = self.materialise(key, recycled)
        self.entries[key] = [value, key.nbytes, 1]
        self.used += key.nbytes
        self.peak = max(self.peak, self.used)
        return value

    def insert(self, key, value):
        """Add a new pinned entry (a result, or None to reserve its bytes);
        returns a recycled value of the same size, or None."""
        recycled = self._make_room(key.nbytes)
        self.entries[key] = [value, key.nbytes, 1]
        self.used += key.nbytes
        self.peak = max(self.peak, self.used)
        return recycled

    def get(self, key):
        return self.entries[key][0]

    def set(self, key, value) -> None:
        self.entries[key][0] = value

    def release(self, key) -> None:
        e = self.entries.get(key)
        if e is not None and e[2] > 0:
            e[2] -= 1


# ---------------------------------------------------------------------------
# GPU side
# ---------------------------------------------------------------------------

def device_propagator(torch, key: PropKey, device, out=None):
    """Synthetic propagator of `key` generated on the device (the stand-in for
    loading a solved propagator into HBM), written into `out` if given."""
    Lx, Ly, Lz, Lt = key.dims
    V = Lx * Ly * Lz * Lt
    cdt = torch.complex128 if key.dtype == "complex128" else torch.complex64
    real = torch.float64 if key.dtype == "complex128" else torch.float32
    if out is None:
        out = torch.empty((V, 4, 4, 3, 3), dtype=cdt, device=device)
    g = torch.Generator(device=device)
    g.manual_seed(key.seed)
    torch.randn((V, 144, 2), dtype=real, device=device, generator=g, out=torch.view_as_real(out).view(V, 144, 2))
    prof = torch.exp(-key.alpha * torch.abs(torch.arange(Lt, dtype=torch.float64, device=device) - key.t0))
    out.view(Lt, -1, 144).mul_(prof.to(real)[:, None, None])
    return out


@dataclasses.dataclass
class Entry:
    tensor: object
    ready: object  # torch.cuda.Event recorded after the tensor was written


@dataclasses.dataclass
class Flight:
    event: object
    requests: List[Request]
    keep: list


class Engine:
    """Continuous-batching executor over one GPU."""

    def __init__(self, cap_bytes: int = 200 << 30, streams: int = 4, max_inflight: int = 8, device="cuda:0",
                 grid_tuning: Optional[Dict[str, Tuple[int, int, int]]] = None, poll_s: float = 5e-5,
                 max_batch: int = 32, batch_baryons: bool = True):
        import torch
        self.torch = torch
        self.device = torch.device(device)
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(streams)]
        self.load_stream = torch.cuda.Stream(device=self.device)
        self.max_inflight = max_inflight
        self.max_batch = max_batch
        self.batch_baryons = batch_baryons
        self.grid_tuning = dict(GRID_TUNING, **(grid_tuning or {}))
        self.poll_s = poll_s
        self.pool = PropagatorPool(cap_bytes, self._materialise)
        self._rr = 0
        self._spare: Dict[int, Entry] = {}
        self.launches = {"meson_batches": 0, "meson_requests": 0, "baryon_batches": 0, "baryon": 0, "diquark": 0}

    # -- pool ----------------------------------------------------------------
    def _materialise(self, key, recycled=None):
        torch = self.torch
        with torch.cuda.stream(self.load_stream):
            cdt = torch.complex128 if key.dtype == "complex128" else torch.complex64
            reuse = self._reuse(recycled, (int(np.prod(key.dims)), 4, 4, 3, 3), cdt, self.load_stream)
            t = device_propagator(torch, key, self.device, out=reuse)
            ev = torch.cuda.Event()
            ev.record(self.load_stream)
        return Entry(t, ev)

    def _reuse(self, recycled, shape, dtype, stream):
        """The memory of an evicted entry of the same byte size as a tensor of
        `shape` and `dtype`, ordered after its last writer; None if there is
        none."""
        if recycled is None or recycled.tensor is None:
            return None
        stream.wait_event(recycled.ready)
        return recycled.tensor.reshape(-1).view(self.torch.uint8).view(dtype).view(shape)

    def _index(self, values):
        """Small index tensor on the device, copied from pinned host memory
        without synchronising the host with the current stream."""
        torch = self.torch
        return torch.tensor(values, dtype=torch.long).pin_memory().to(self.device, non_blocking=True)

    def _next_stream(self):
        s = self.streams[self._rr % len(self.streams)]
        self._rr += 1
        return s

    def _inputs(self, r: Request, stream):
        ts = []
        for k in r.props:
            e = self.pool.get(k)
            stream.wait_event(e.ready)
            ts.append(e.tensor)
        return ts

    # -- launches --------------------------------------------------------------
    @staticmethod
    def _momentum_groups(union: Sequence[Tuple[int, int, int]]) -> List[List[Tuple[int, int, int]]]:
        """The union of momenta cut into groups of at most 16 slots, a slot
        being a momentum together with its negative when both are wanted (the
        grid kernel accumulates k and -k as one cos/sin pair), so that one grid
        pass (16 slots) reads every propagator once for up to 32 momenta: the
        27 momenta of {-1,0,1}^3 take one pass."""
        have = set(union)
        slots, seen = [], set()
        for m in union:
            if m in seen:
                continue
            neg = (-m[0], -m[1], -m[2])
            slot = [m] if neg == m or neg not in have else [m, neg]
            seen.update(slot)
            slots.append(slot)
        return [[m for s in slots[g:g + 16] for m in s] for g in range(0, len(slots), 16)]

    def _launch_mesons(self, reqs: List[Request], stream) -> list:
        """One batched grid launch per group of 16 momentum slots of the union
        of the requests' momenta, each over the requests that need one of
        them; then every request gathers its (gamma pair, momentum) entries."""
        torch = self.torch
        keep = []
        union = sorted({tuple(m) for r in reqs for m in r.momenta})
        for r in reqs:
            r.result = torch.empty((len(r.gammas), len(r.momenta), r.dims[3]), dtype=torch.complex128,
                                   device=self.device)
        by_dtype: Dict[str, List[Request]] = {}
        for r in reqs:
            by_dtype.setdefault(r.props[0].dtype, []).append(r)
        with torch.cuda.stream(stream):
            for dt, group in by_dtype.items():
                tuning = self.grid_tuning[dt]
                inputs = {r.rid: self._inputs(r, stream) for r in group}
                for moms in self._momentum_groups(union):
                    sub = [r for r in group if any(tuple(m) in moms for m in r.momenta)]
                    if not sub:
                        continue
                    batch = [(inputs[r.rid][0], inputs[r.rid][1], r.dims, 0) for r in sub]
                    ws = torch.empty(pc.meson_grid_workspace_bytes(batch, len(moms), tuning), dtype=torch.uint8,
                                     device=self.device)
                    outs = pc.mesonGridProject(batch, moms, tuning=tuning, workspace=ws, stream=stream.cuda_stream)
                    keep += [ws] + outs
                    for r, o in zip(sub, outs):
                        # this pass's momenta of r: slot k in the pass -> position in r's result
                        pos = [(moms.index(tuple(m)), pi) for pi, m in enumerate(r.momenta) if tuple(m) in moms]
                        pairs = self._index([a * 16 + b for a, b in r.gammas])
                        ks = self._index([k for k, _ in pos])
                        pis = self._index([pi for _, pi in pos])
                        picked = o.reshape(256, len(moms), r.dims[3]).index_select(0, pairs).index_select(1, ks)
                        r.result.index_copy_(1, pis, picked)
                        keep += [pairs, ks, pis, picked]
        self.launches["meson_batches"] += 1
        self.launches["meson_requests"] += len(reqs)
        return keep

    def _launch_baryons(self, reqs: List[Request], stream) -> list:
        """All baryon requests of a tick in batched fused launches
        (contract_baryon_twopoint_project_batch, up to 16 per launch) over the
        union of their insertions and momenta; every request then gathers
        its own (insertion, momentum) entries.  Momenta of {-1, 0, 1}^3 pair
        into at most 14 +-k slots, one MFMA tile, so the union costs the
        kernel nothing."""
        torch = self.torch
        keep = []
        gammas = sorted({int(g) for r in reqs for g in r.gammas})
        moms = sorted({tuple(m) for r in reqs for m in r.momenta})
        by_dtype: Dict[str, List[Request]] = {}
        for r in reqs:
            by_dtype.setdefault(r.props[0].dtype, []).append(r)
        with torch.cuda.stream(stream):
            for dt, group in by_dtype.items():
                batch = [(*self._inputs(r, stream), r.dims) for r in group]
                ws = torch.empty(pc.baryon_batch_workspace_bytes(batch, len(moms)), dtype=torch.uint8,
                                 device=self.device)
                outs = pc.baryonTwoPointProjectBatch(batch, PROTON, CG5, CG5, gammas, moms, workspace=ws,
                                                     stream=stream.cuda_stream)
                keep += [ws] + outs
                for r, o in zip(group, outs):
                    gi = self._index([gammas.index(int(g)) for g in r.gammas])
                    mi = self._index([moms.index(tuple(m)) for m in r.momenta])
                    r.result = o.index_select(0, gi).index_select(1, mi)
                    keep += [gi, mi]
        self.launches["baryon_batches"] += 1
        self.launches["baryon"] += len(reqs)
        return keep

    def _launch_diquark(self, r: Request, stream) -> list:
        torch = self.torch
        key = r.out_key()
        with torch.cuda.stream(stream):
            pi, pj = self._inputs(r, stream)
            out = self._reuse(self._spare.pop(r.rid, None), pi.shape, pi.dtype, stream)
            if out is None:
                # pool memory is allocated on the load stream only: torch's
                # caching allocator reuses a freed block only for allocations
                # on the stream it was allocated on, so evicted entries'
                # memory then serves the next entries (a fresh multi-GB
                # allocation costs the host tens of milliseconds).  Entries
                # are evicted only after their last reader has finished.
                with torch.cuda.stream(self.load_stream):
                    out = torch.empty_like(pi)
            (gij, gkl), = r.gammas[:1]
            with pc.abi_stream(stream):
                pc.baryon_diquark(out, pi, pj, int(np.prod(r.dims)), int(gij), int(gkl))
            ev = torch.cuda.Event()
            ev.record(stream)
        self.pool.set(key, Entry(out, ev))  # its bytes were reserved at admission
        r.result = key
        self.launches["diquark"] += 1
        return []

    # -- scheduler -------------------------------------------------------------
    def _admissible(self, r: Request) -> bool:
        extra = [r.out_key()] if r.kind == "diquark" else []
        return self.pool.fits(r.props, extra)

    def _complete(self, f: Flight, now: float) -> None:
        for r in f.requests:
            r.done = now
            for k in r.props:
                self.pool.release(k)
            if r.kind == "diquark":
                self.pool.release(r.out_key())

    def run(self, requests: Sequence[Request], realtime: bool = True) -> Dict[str, object]:
        """Serve `requests` (sorted by arrival time in seconds from the start).
        realtime=False admits everything at once (arrival times ignored for
        admission but kept for latency)."""
        torch = self.torch
        for r in requests:
            need = sum(k.nbytes for k in set(r.props)) + (r.out_key().nbytes if r.kind == "diquark" else 0)
            if need > self.pool.cap:
                raise PoolFull(f"request {r.rid} needs {need} bytes, more than the pool cap {self.pool.cap}")
        if not realtime:
            for r in requests:
                r.arrival = 0.0
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        queue: deque = deque()
        flights: List[Flight] = []
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        batch_sizes = []
        while pending or queue or flights:
            now = time.perf_counter() - t_start
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            progressed = False
            alive = []
            for f in flights:
                if f.event.query():
                    self._complete(f, time.perf_counter() - t_start)
                    progressed = True
                else:
                    alive.append(f)
            flights = alive
            admitted: List[Request] = []
            slots = self.max_inflight - len(flights)
            nbatch = {"meson": 0, "baryon": 0}
            while queue and slots > 0:
                r = queue[0]
                if (r.kind in nbatch and nbatch[r.kind] == self.max_batch) or not self._admissible(r):
                    break
                queue.popleft()
                for k in r.props:
                    self.pool.acquire(k)
                if r.kind == "diquark":  # reserve the output's bytes
                    spare = self.pool.insert(r.out_key(), None)
                    if spare is not None:
                        self._spare[r.rid] = spare
                r.admitted = time.perf_counter() - t_start
                admitted.append(r)
                if r.kind == "diquark" or (r.kind == "baryon" and not self.batch_baryons):
                    slots -= 1
                elif r.kind == "meson":  # one batched launch for all mesons of this tick
                    nbatch["meson"] += 1
                    slots -= nbatch["meson"] == 1
                else:  # one batched launch per lattice for the baryons of this tick
                    nbatch["baryon"] += 1
                    slots -= not any(a.kind == "baryon" and a.dims == r.dims for a in admitted[:-1])
            if admitted:
                progressed = True
                # mesons: one flight for the tick; baryons: one flight per
                # lattice, so a batch on a small lattice completes (and
                # releases its pinned propagators) without waiting for a large one
                groups = [("meson", [r for r in admitted if r.kind == "meson"])]
                if self.batch_baryons:
                    for dims in sorted({r.dims for r in admitted if r.kind == "baryon"}):
                        groups.append(("baryon", [r for r in admitted if r.kind == "baryon" and r.dims == dims]))
                else:
                    groups += [("baryon", [r]) for r in admitted if r.kind == "baryon"]
                for kind, group in groups:
                    if group:
                        s = self._next_stream()
                        keep = (self._launch_mesons if kind == "meson" else self._launch_baryons)(group, s)
                        ev = torch.cuda.Event()
                        ev.record(s)
                        flights.append(Flight(ev, group, keep))
                        if kind == "meson":
                            batch_sizes.append(len(group))
                for r in admitted:
                    if r.kind != "diquark":
                        continue
                    s = self._next_stream()
                    keep = self._launch_diquark(r, s)
                    ev = torch.cuda.Event()
                    ev.record(s)
                    flights.append(Flight(ev, [r], keep))
            if not progressed:
                wait = self.poll_s
                if not queue and not flights and pending:
                    wait = max(0.0, min(pending[0].arrival - now, 0.01))
                time.sleep(wait)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_start
        return summarise(requests, wall, self.pool, batch_sizes, self.launches)


def percentile(xs: Sequence[float], q: float) -> float:
    return float(np.percentile(np.asarray(xs), q)) if len(xs) else float("nan")


def summarise(requests, wall, pool, batch_sizes, launches) -> Dict[str, object]:
    lat = [r.latency for r in requests if r.latency is not None]
    per_kind = {}
    for k in KINDS:
        lk = [r.latency for r in requests if r.kind == k and r.latency is not None]
        if lk:
            per_kind[k] = {"n": len(lk), "p50_ms": 1e3 * percentile(lk, 50), "p99_ms": 1e3 * percentile(lk, 99)}
    span = max((r.done for r in requests if r.done is not None), default=wall)
    return {
        "requests": len(requests),
        "completed": len(lat),
        "wall_s": wall,
        "throughput_rps": len(lat) / span if span > 0 else float("nan"),
        "p50_ms": 1e3 * percentile(lat, 50),
        "p99_ms": 1e3 * percentile(lat, 99),
        "mean_ms": 1e3 * float(np.mean(lat)) if lat else float("nan"),
        "per_kind": per_kind,
        "pool": {"cap_GB": pool.cap / 1e9, "peak_GB": pool.peak / 1e9, "hits": pool.hits, "misses": pool.misses,
                 "evictions": pool.evictions},
        "meson_batch_mean": float(np.mean(batch_sizes)) if batch_sizes else 0.0,
        "launches": dict(launches),
    }


# ---------------------------------------------------------------------------
# Synthetic request stream
# ---------------------------------------------------------------------------

LATTICES = [(16, 16, 16, 32), (24, 24, 24, 64), (32, 32, 32, 64), (48, 48, 48, 96)]


def make_requests(rate: float, duration: float, seed: int, lattices=LATTICES, catalogue: int = 4,
                  mix=(0.5, 0.3, 0.2), max_gammas: int = 16, max_momenta: int = 27, alpha: float = 0.25,
                  first_rid: int = 0, dtype: str = "complex128") -> List[Request]:
    """Poisson arrivals at `rate` per second over `duration` seconds: 50% meson,
    30% baryon, 20% diquark requests on a lattice drawn from `lattices`, each
    referencing propagators of a per-lattice catalogue of `catalogue` entries,
    with 1..max_gammas gammas and 1..max_momenta momenta from {-1,0,1}^3."""
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    rid = first_rid
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        kind = KINDS[int(rng.choice(3, p=mix))]
        dims = tuple(lattices[int(rng.integers(len(lattices)))])
        keys = [PropKey(dims, i, t0=(7 * i) % dims[3], alpha=alpha, dtype=dtype) for i in range(catalogue)]
        ng = int(rng.integers(1, max_gammas + 1))
        nm = int(rng.integers(1, max_momenta + 1))
        moms = [MOM27[i] for i in sorted(rng.choice(27, size=min(nm, 27), replace=False))]
        if kind == "meson":
            a, b = rng.choice(catalogue, size=2, replace=catalogue < 2)
            props = (keys[a], keys[b])
            gammas = [(int(rng.integers(16)), int(rng.integers(16))) for _ in range(ng)]
        elif kind == "baryon":
            a, b = rng.choice(catalogue, size=2, replace=catalogue < 2)
            props = (keys[a], keys[b], keys[a])
            gammas = [int(g) for g in rng.choice(16, size=ng, replace=False)]
        else:
            a, b = rng.choice(catalogue, size=2, replace=catalogue < 2)
            props = (keys[a], keys[b])
            gammas = [(int(rng.integers(16)), int(rng.integers(16)))]
            moms = []
        out.append(Request(rid, kind, dims, props, gammas, moms, arrival=t))
        rid += 1
    return out


def sweep(rates: Sequence[float], duration: float, seed: int = 7, cap_gb: float = 200.0, streams: int = 4,
          max_inflight: int = 8, lattices=LATTICES, catalogue: int = 4, warm: bool = True,
          batch_baryons: bool = True) -> Dict[str, object]:
    """Serve one Poisson stream per rate on a shared engine (the pool stays warm
    between rates) and report throughput and latency per rate."""
    eng = Engine(cap_bytes=int(cap_gb * 1e9), streams=streams, max_inflight=max_inflight,
                 batch_baryons=batch_baryons)
    if warm:  # load the catalogue, largest lattices last (most recently used)
        for dims in lattices:
            for i in range(catalogue):
                k = PropKey(tuple(dims), i, t0=(7 * i) % dims[3])
                if eng.pool.fits([k]):
                    eng.pool.acquire(k)
                    eng.pool.release(k)
        eng.torch.cuda.synchronize()
    res = {"rates": {}}
    rid = 0
    for k, rate in enumerate(rates):
        reqs = make_requests(rate, duration, seed + k, lattices=lattices, catalogue=catalogue, first_rid=rid)
        rid += len(reqs)
        print(f"serving: rate {rate} req/s, {len(reqs)} requests", file=sys.stderr, flush=True)
        res["rates"][str(rate)] = eng.run(reqs)
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--rates", default="10,25,50,100,200")
    ap.add_argument("--duration", type=float, default=3.0)
    ap.add_argument("--cap-gb", type=float, default=200.0)
    ap.add_argument("--streams", type=int, default=4)
    ap.add_argument("--max-inflight", type=int, default=8)
    ap.add_argument("--seed", type=int, default=7)
    args = ap.parse_args(argv)
    rates = [float(r) for r in args.rates.split(",")]
    print(json.dumps(sweep(rates, args.duration, args.seed, args.cap_gb, args.streams, args.max_inflight)))


if __name__ == "__main__":
    main()
