"""The autotuner on the GPU: it times every candidate of a small candidate
set on the caller's batch, the winner is never slower than the library
default in the same table, results under the winner match the FP64 oracle
(1e-12 per timeslice), and a second tuning of the same batch signature comes
from the cache file without timing anything."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

MOM7 = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
PROTON = [(0, 1.0), (5, 1.0)]


def per_timeslice_error(got, want):
    num = np.abs(got - want).reshape(-1, got.shape[-1]).max(axis=0)
    den = np.abs(want).reshape(-1, want.shape[-1]).max(axis=0)
    return float((num / np.maximum(den, 1e-300)).max())


@pytest.fixture
def tuner(cuda, tmp_path, monkeypatch):
    import torch
    from contract_amd import autotune
    from contract_amd import pycontract as pc
    monkeypatch.setenv("CONTRACT_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    autotune.clear_cache()
    yield torch, pc, autotune, cuda
    autotune.clear_cache()


def test_meson_grid_tuning(tuner):
    torch, pc, autotune, cuda = tuner
    lats = [(8, 8, 8, 8), (4, 4, 4, 16)]
    host = [(oracle.synthetic_propagator(d, 10 + k, 0.3), oracle.synthetic_propagator(d, 20 + k, 0.3)) for k, d in
            enumerate(lats)]
    reqs = [(torch.from_numpy(a).to(cuda), torch.from_numpy(b).to(cuda), d, 0) for (a, b), d in zip(host, lats)]
    default = (1024, 4, 2, 2)
    cands = autotune.candidates(sites=(256, 1024), tiles=(4, 8), blocks=(1, 2), units=(1, 2))
    assert default in cands
    best, table = autotune.tune_meson_grid(reqs, MOM7, cands=cands, reps=2)
    assert set(table) == set(cands) and best in cands
    assert table[best] <= table[default]
    outs = pc.mesonGridProject(reqs, MOM7, tuning=best)
    for (a, b), d, o in zip(host, lats, outs):
        assert per_timeslice_error(o.cpu().numpy(), oracle.meson_grid_project(a, b, d, MOM7)) < 1e-12
    again, t2 = autotune.tune_meson_grid(reqs, MOM7, cands=cands)
    assert again == best and list(t2) == [best]
    autotune.clear_cache()  # memory only: the file answers
    again, t3 = autotune.tune_meson_grid(reqs, MOM7, cands=cands)
    assert again == best and list(t3) == [best]


def test_meson_grid_tuning_plain_fp32(tuner):
    """Plain complex64 storage: the tuner times the vector-ALU and the
    wave-local forms (no shared-tile matrix-core form) and the winner's
    results are within 1e-6 of the oracle on the same complex64 values."""
    torch, pc, autotune, cuda = tuner
    lats = [(8, 8, 8, 8), (4, 4, 4, 16)]
    host = [(oracle.synthetic_propagator(d, 40 + k, 0.3).astype(np.complex64),
             oracle.synthetic_propagator(d, 50 + k, 0.3).astype(np.complex64)) for k, d in enumerate(lats)]
    reqs = [(torch.from_numpy(a).to(cuda), torch.from_numpy(b).to(cuda), d, 0) for (a, b), d in zip(host, lats)]
    units = autotune._units_allowed(reqs)
    assert set(units) == {1, 3}
    cands = autotune.candidates(sites=(256, 1024), tiles=(4, 8), blocks=(2,), units=units)
    best, table = autotune.tune_meson_grid(reqs, MOM7, cands=cands, reps=2)
    assert set(table) == set(cands) and {u for *_, u in table} == {1, 3}
    outs = pc.mesonGridProject(reqs, MOM7, tuning=best)
    for (a, b), d, o in zip(host, lats, outs):
        want = oracle.meson_grid_project(a.astype(np.complex128), b.astype(np.complex128), d, MOM7)
        assert per_timeslice_error(o.cpu().numpy(), want) < 1e-6


def test_baryon_plan_tuning(tuner):
    torch, pc, autotune, cuda = tuner
    dims = (8, 8, 8, 16)
    ps = [oracle.synthetic_propagator(dims, s, 0.3) for s in (31, 32)]
    u, d = (torch.from_numpy(p).to(cuda) for p in ps)
    reqs = [(u, d, u, dims)]
    cands = autotune.baryon_candidates(items=(4, 8, 16), sites=(64, 128))
    best, table = autotune.tune_baryon(reqs, PROTON, 5, 5, list(range(16)), MOM7, cands=cands, reps=3)
    assert set(table) == set(cands) and table[best] <= table[autotune.BARYON_DEFAULT]
    with pc.baryon_plan(best):
        got = pc.baryonTwoPointProjectBatch(reqs, PROTON, 5, 5, list(range(16)), MOM7)[0].cpu().numpy()
    assert pc.get_baryon_plan() == autotune.BARYON_DEFAULT
    want = oracle.baryon_twopoint_project(ps[0], ps[1], ps[0], PROTON, 5, 5, list(range(16)), dims, MOM7)
    assert per_timeslice_error(got, want) < 1e-12
    again, t2 = autotune.tune_baryon(reqs, PROTON, 5, 5, list(range(16)), MOM7, cands=cands)
    assert again == best and list(t2) == [best]
