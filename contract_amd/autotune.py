"""Autotuner for the fused kernels' work decomposition.

Two launch families have knobs whose best setting depends on the batch:

* the batched meson Gamma x Gamma' grid (contract_meson_grid_project,
  contract_meson_grid_tuning): sites per work item (how finely timeslices are
  cut, i.e. load balance against the 64 KiB partial each item writes), the
  staged tile (4 or 8 sites per LDS step: fewer barriers against more
  registers and LDS), persistent workgroups per compute unit (occupancy
  against LDS/register pressure) and the unit that forms the per-site traces
  (vector ALUs, or matrix cores for FP64 and site-scaled FP32 storage);
* the fused baryon two-point (contract_baryon_twopoint_project*, set through
  contract_set_baryon_plan): work items per compute unit and the minimum
  sites per work item.

The tuner times every candidate on the caller's own batch, on the caller's
stream, and keeps the winner per (device, batch signature): in memory for the
life of the process and in a JSON cache file (``CONTRACT_AMD_TUNE_CACHE``, by
default ``~/.cache/contract_amd/autotune.json``), so a later process on the
same device starts from it.
"""

from __future__ import annotations

import itertools
import json
import os
import tempfile
from typing import Dict, List, Optional, Sequence, Tuple

from . import pycontract as pc

DEFAULT_SITES = (256, 512, 1024, 2048, 4096)
DEFAULT_TILES = (4, 8)
DEFAULT_BLOCKS = (1, 2, 3, 4)
DEFAULT_UNITS = (1, 2, 3)  # trace_unit: vector ALUs, matrix cores (shared tile), matrix cores (wave-local K)
WAVE_LOCAL_BLOCKS = (4, 8)  # trace_unit 3: one-wave workgroups, one or two per SIMD (tile fixed at 4 sites)

BARYON_ITEMS_PER_CU = (4, 6, 8, 12, 16)
BARYON_MIN_SITES = (64, 128, 256, 512)
BARYON_DEFAULT = (8, 128)

GRID_SLOTS_PER_PASS = 16  # momentum slots per grid pass (include/contract_ext.h)
GRID_SLOTS_PER_TILE = 8   # slots per MFMA row tile

_cache: Dict[tuple, Tuple[tuple, float]] = {}


# ---------------------------------------------------------------------------
# signatures
# ---------------------------------------------------------------------------

def momentum_slots(momenta) -> int:
    """Number of +-k slots the library forms from `momenta`: a momentum and
    its negative share a slot (each listed momentum pairs with at most one
    later listed negative, the zero momentum with none), as the library's
    pairing does."""
    moms = [tuple(int(v) for v in m) for m in momenta]
    used = [False] * len(moms)
    n = 0
    for p, m in enumerate(moms):
        if used[p]:
            continue
        used[p] = True
        n += 1
        if m == (0, 0, 0):
            continue
        neg = (-m[0], -m[1], -m[2])
        for q in range(p + 1, len(moms)):
            if not used[q] and moms[q] == neg:
                used[q] = True
                break
    return n


def grid_passes(momenta) -> Tuple[int, ...]:
    """Row tiles of every grid pass for these momenta: passes of up to 16
    slots, one MFMA row tile per 8 slots.  Batches with the same passes run
    the same kernels."""
    n = momentum_slots(momenta)
    return tuple((min(GRID_SLOTS_PER_PASS, n - p0) + GRID_SLOTS_PER_TILE - 1) // GRID_SLOTS_PER_TILE
                 for p0 in range(0, n, GRID_SLOTS_PER_PASS))


def _storage(p) -> str:
    return "fp32_scaled" if isinstance(p, pc.ScaledPropagator) else str(p.dtype)


def signature(requests, momenta) -> tuple:
    """What the best grid tuning depends on: the storage form, the grid
    passes (kernel form and count, from the +-k slots of the momenta) and the
    lattices in batch order (the item space is laid out request by
    request)."""
    lat = tuple((tuple(int(v) for v in r[2]), int(r[3]) if len(r) > 3 else 0) for r in requests)
    return ("meson_grid", _storage(requests[0][0]), grid_passes(momenta), lat)


def baryon_signature(requests, terms, gamma_ij: int, momenta, nproj: int) -> tuple:
    """What the best baryon plan depends on: the storage form, whether P_n is
    P_i (the shared-propagator kernel), the compile-time term and diquark
    forms, the momentum tiles (16 slots each) and the lattices."""
    first = requests[0]
    share = all(r[2] is r[0] or pc._ptr(r[2]) == pc._ptr(r[0]) for r in requests)
    nucleon = [int(t) for t, _ in terms] == [0, 5]
    lat = tuple((tuple(int(v) for v in r[3]), int(r[4]) if len(r) > 4 else 0) for r in requests)
    tiles = (momentum_slots(momenta) + 15) // 16
    return ("baryon", _storage(first[0]), share, nucleon, int(gamma_ij) == 5, tiles, lat)


# ---------------------------------------------------------------------------
# cache
# ---------------------------------------------------------------------------

def cache_path() -> str:
    return os.environ.get("CONTRACT_AMD_TUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "contract_amd", "autotune.json")


def _device_name(device) -> str:
    import torch
    try:
        return torch.cuda.get_device_name(device)
    except Exception:  # noqa: BLE001 - the name only keys the cache
        return "unknown"


def _file_key(device: str, key: tuple) -> str:
    return json.dumps([device, key], separators=(",", ":"))


def load_cached(device: str, key: tuple) -> Optional[Tuple[tuple, float]]:
    """The cache file's entry for (device, key), or None."""
    try:
        with open(cache_path()) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return None
    e = data.get(_file_key(device, key)) if isinstance(data, dict) else None
    if not isinstance(e, dict) or "best" not in e:
        return None
    return tuple(int(v) for v in e["best"]), float(e.get("ms", float("nan")))


def store_cached(device: str, key: tuple, best: tuple, ms: float) -> None:
    """Add (device, key) -> best to the cache file (written atomically; a
    cache that cannot be written is skipped)."""
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        try:
            with open(path) as f:
                data = json.load(f)
            if not isinstance(data, dict):
                data = {}
        except (OSError, ValueError):
            data = {}
        data[_file_key(device, key)] = {"best": list(best), "ms": ms}
        fd, tmp = tempfile.mkstemp(dir=os.path.dirname(path), prefix=".autotune")
        with os.fdopen(fd, "w") as f:
            json.dump(data, f, indent=0, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def _lookup(device: str, key: tuple, use_cache: bool):
    if not use_cache:
        return None
    if (device, key) in _cache:
        return _cache[(device, key)]
    hit = load_cached(device, key)
    if hit is not None:
        _cache[(device, key)] = hit
    return hit


def clear_cache(file: bool = False) -> None:
    """Forget the in-process winners (and the cache file with file=True)."""
    _cache.clear()
    if file:
        try:
            os.remove(cache_path())
        except OSError:
            pass


# ---------------------------------------------------------------------------
# meson grid
# ---------------------------------------------------------------------------

def candidates(sites: Sequence[int] = DEFAULT_SITES, tiles: Sequence[int] = DEFAULT_TILES,
               blocks: Sequence[int] = DEFAULT_BLOCKS, units: Sequence[int] = DEFAULT_UNITS
               ) -> List[Tuple[int, int, int, int]]:
    """(sites_per_item, tile_sites, blocks_per_cu, trace_unit) plans.  The
    wave-local form (unit 3) stages 4-site tiles in one-wave workgroups, one
    per SIMD: only its sites per item vary."""
    out = [(s, t, b, u) for s, t, b, u in itertools.product(sites, tiles, blocks, [u for u in units if u != 3])]
    if 3 in units:
        out += [(s, 4, b, 3) for s in sites for b in WAVE_LOCAL_BLOCKS]
    return out


def _units_allowed(requests) -> Sequence[int]:
    """Plain complex64 storage has no shared-tile matrix-core form (its
    wave-local form scales each site as the tile lands)."""
    import torch
    return (1, 3) if getattr(requests[0][0], "dtype", None) == torch.complex64 else DEFAULT_UNITS


def _time(fn, stream, reps: int, warmup: int) -> float:
    import torch
    for _ in range(warmup):
        fn()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        fn()
    stop.record(stream)
    stop.synchronize()
    return start.elapsed_time(stop) / reps


def tune_meson_grid(requests, momenta, cands=None, reps: int = 3, warmup: int = 1, workspace=None,
                    use_cache: bool = True):
    """Time every candidate tuning on this batch and return
    (best (sites_per_item, tile_sites, blocks_per_cu, trace_unit), {tuning: ms}).
    A cached winner returns at once as (best, {best: ms}).

    ``workspace`` (optional) must be large enough for every candidate; the
    tuner otherwise allocates the largest one once.  Outputs are written to
    scratch tensors, so the caller's results are untouched."""
    import torch
    nmom = len(momenta)
    key = signature(requests, momenta)
    dev = requests[0][0].device
    name = _device_name(dev)
    hit = _lookup(name, key, use_cache)
    if hit is not None:
        return hit[0], {hit[0]: hit[1]}
    cands = list(cands) if cands is not None else candidates(units=_units_allowed(requests))
    need = max(pc.meson_grid_workspace_bytes(requests, nmom, c) for c in cands)
    if workspace is None or workspace.numel() * workspace.element_size() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device=dev)
    outs = [torch.empty((16, 16, nmom, int(r[2][3])), dtype=torch.complex128, device=dev) for r in requests]
    stream = torch.cuda.current_stream()
    table = {}
    for c in cands:
        table[c] = _time(lambda: pc.mesonGridProject(requests, momenta, tuning=c, outs=outs, workspace=workspace,
                                                     stream=stream.cuda_stream), stream, reps, warmup)
    best = min(table, key=table.get)
    _cache[(name, key)] = (best, table[best])
    if use_cache:
        store_cached(name, key, best, table[best])
    return best, table


# ---------------------------------------------------------------------------
# fused baryon two-point
# ---------------------------------------------------------------------------

def baryon_candidates(items: Sequence[int] = BARYON_ITEMS_PER_CU, sites: Sequence[int] = BARYON_MIN_SITES
                      ) -> List[Tuple[int, int]]:
    c = [(i, s) for i, s in itertools.product(items, sites)]
    if BARYON_DEFAULT not in c:
        c.append(BARYON_DEFAULT)
    return c


def tune_baryon(requests, terms, gamma_ij: int, gamma_kl: int, insertions, momenta, cands=None, reps: int = 5,
                warmup: int = 1, use_cache: bool = True):
    """Time every (items_per_cu, min_item_sites) plan on this baryon batch
    (tuples (Pi, Pj, Pn, dims[, site_order...]) as for
    baryonTwoPointProjectBatch) and return (best plan, {plan: ms}); the
    library default (8, 128) is always among the candidates.  Use the winner
    with ``pycontract.baryon_plan(best)``."""
    import torch
    key = baryon_signature(requests, terms, gamma_ij, momenta, len(insertions))
    dev = requests[0][0].device
    name = _device_name(dev)
    hit = _lookup(name, key, use_cache)
    if hit is not None:
        return hit[0], {hit[0]: hit[1]}
    cands = list(cands) if cands is not None else baryon_candidates()
    nmom = len(momenta)
    outs = [torch.empty((len(insertions), nmom, int(r[3][3])), dtype=torch.complex128, device=dev)
            for r in requests]
    stream = torch.cuda.current_stream()
    table = {}
    for c in cands:
        with pc.baryon_plan(c):
            ws = torch.empty(max(pc.baryon_batch_workspace_bytes(requests, nmom), 1), dtype=torch.uint8,
                             device=dev)
            table[c] = _time(lambda: pc.baryonTwoPointProjectBatch(
                requests, terms, gamma_ij, gamma_kl, insertions, momenta, outs=outs, workspace=ws,
                stream=stream.cuda_stream), stream, reps, warmup)
        del ws
    best = min(table, key=table.get)
    _cache[(name, key)] = (best, table[best])
    if use_cache:
        store_cached(name, key, best, table[best])
    return best, table
