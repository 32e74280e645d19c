"""Autotuner for the batched meson Gamma x Gamma' grid
(contract_meson_grid_project).

The kernel's work decomposition has four knobs (contract_meson_grid_tuning):
sites per work item (how finely timeslices are cut, i.e. load balance against
the 64 KiB partial each item writes), the staged tile (4 or 8 sites per LDS
step: fewer barriers against more registers and LDS), persistent workgroups
per compute unit (occupancy against LDS/register pressure) and the unit that
forms the per-site traces (vector ALUs, or matrix cores for FP64 and
site-scaled FP32 storage).  The
best choice depends on the batch (lattice mix, precision), so the tuner times
every candidate on the caller's own batch, on the caller's stream, and caches
the winner per batch signature for the life of the process.
"""

from __future__ import annotations

import itertools
from typing import Dict, List, Sequence, Tuple

from . import pycontract as pc

DEFAULT_SITES = (256, 512, 1024, 2048, 4096)
DEFAULT_TILES = (4, 8)
DEFAULT_BLOCKS = (1, 2, 3)
DEFAULT_UNITS = (1, 2)  # trace_unit: vector ALUs, matrix cores

_cache: Dict[tuple, Tuple[Tuple[int, int, int], float]] = {}


def signature(requests, nmom: int) -> tuple:
    """What the best tuning depends on: dtype, the kernel form (one MFMA row
    tile per pass up to 8 momenta, two beyond: a momentum and its negative
    share a slot, 8 slots a tile), the number of passes (at most 32 momenta
    each) and the lattices in batch order (the item space is laid out request
    by request)."""
    dt = str(requests[0][0].dtype)
    lat = tuple((tuple(int(v) for v in r[2]), int(r[3]) if len(r) > 3 else 0) for r in requests)
    return (dt, 1 if nmom <= 8 else 2, (nmom + 31) // 32, lat)


def candidates(sites: Sequence[int] = DEFAULT_SITES, tiles: Sequence[int] = DEFAULT_TILES,
               blocks: Sequence[int] = DEFAULT_BLOCKS, units: Sequence[int] = DEFAULT_UNITS
               ) -> List[Tuple[int, int, int, int]]:
    return [(s, t, b, u) for s, t, b, u in itertools.product(sites, tiles, blocks, units)]


def _units_allowed(requests) -> Sequence[int]:
    """Plain complex64 storage has no matrix-core trace form."""
    import torch
    return (1,) if getattr(requests[0][0], "dtype", None) == torch.complex64 else DEFAULT_UNITS


def tune_meson_grid(requests, momenta, cands=None, reps: int = 3, warmup: int = 1, workspace=None,
                    use_cache: bool = True):
    """Time every candidate tuning on this batch and return
    (best (sites_per_item, tile_sites, blocks_per_cu, trace_unit), {tuning: ms}).

    ``workspace`` (optional) must be large enough for every candidate; the
    tuner otherwise allocates the largest one once.  Outputs are written to
    scratch tensors, so the caller's results are untouched."""
    import torch
    nmom = len(momenta)
    key = signature(requests, nmom)
    if use_cache and key in _cache:
        best, ms = _cache[key]
        return best, {best: ms}
    cands = list(cands) if cands is not None else candidates(units=_units_allowed(requests))
    need = max(pc.meson_grid_workspace_bytes(requests, nmom, c) for c in cands)
    dev = requests[0][0].device
    if workspace is None or workspace.numel() * workspace.element_size() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device=dev)
    outs = [torch.empty((16, 16, nmom, int(r[2][3])), dtype=torch.complex128, device=dev) for r in requests]
    stream = torch.cuda.current_stream()
    table = {}
    for c in cands:
        for _ in range(warmup):
            pc.mesonGridProject(requests, momenta, tuning=c, outs=outs, workspace=workspace,
                                stream=stream.cuda_stream)
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(reps):
            pc.mesonGridProject(requests, momenta, tuning=c, outs=outs, workspace=workspace,
                                stream=stream.cuda_stream)
        stop.record(stream)
        stop.synchronize()
        table[c] = start.elapsed_time(stop) / reps
    best = min(table, key=table.get)
    _cache[key] = (best, table[best])
    return best, table


def clear_cache() -> None:
    _cache.clear()
