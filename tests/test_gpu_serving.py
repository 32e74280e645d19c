"""The serving loop on the GPU: a mixed stream served with continuous batching
(meson requests batched into shared launches, everything spread over several
streams, pool evictions under a tight cap) gives every request the same result
as serving it alone, and those results match the FP64 oracle.

Tolerance: alone vs in the mix, 1e-12 relative per timeslice (the kernels and
work decompositions are the same, so the results are expected to be
bit-identical); against the oracle, 1e-12 as in the kernel parity tests."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

LATTICES = [(4, 4, 4, 8), (6, 4, 2, 6), (8, 4, 4, 4)]


@pytest.fixture(scope="module")
def srv(cuda):
    import torch
    from contract_amd import serving
    return torch, serving, cuda


def snapshot(torch, eng, r):
    if r.kind == "diquark":
        return eng.pool.get(r.result).tensor.cpu().numpy().copy()
    return r.result.cpu().numpy().copy()


def rel_err(got, want):
    if got.ndim == 3:  # [gamma][momentum][t]
        num = np.abs(got - want).max(axis=(0, 1))
        den = np.abs(want).max(axis=(0, 1))
        return float((num / np.maximum(den, 1e-300)).max())
    return float(np.abs(got - want).max() / np.abs(want).max())


def serve(torch, serving, reqs, cap_bytes, streams=4, max_inflight=8):
    eng = serving.Engine(cap_bytes=cap_bytes, streams=streams, max_inflight=max_inflight)
    eng.run(reqs, realtime=False)
    return eng


def host_prop(torch, serving, key, cuda):
    return serving.device_propagator(torch, key, cuda).cpu().numpy()


def oracle_result(torch, serving, r, cuda):
    ps = [host_prop(torch, serving, k, cuda) for k in r.props]
    if r.kind == "meson":
        grid = oracle.meson_grid_project(ps[0], ps[1], r.dims, r.momenta)
        return np.stack([grid[a, b] for a, b in r.gammas])
    if r.kind == "baryon":
        return oracle.baryon_twopoint_project(*ps, serving.PROTON, serving.CG5, serving.CG5, r.gammas, r.dims,
                                              r.momenta)
    (gij, gkl), = r.gammas
    return oracle.baryon_diquark(ps[0], ps[1], gij, gkl)


def test_mixed_stream_equals_alone_and_oracle(srv):
    torch, serving, cuda = srv
    reqs = serving.make_requests(3000.0, 0.02, seed=11, lattices=LATTICES, catalogue=3, max_gammas=6,
                                 max_momenta=12)
    assert len(reqs) >= 30 and {r.kind for r in reqs} == set(serving.KINDS)
    eng = serve(torch, serving, reqs, cap_bytes=1 << 34)
    assert all(r.done is not None for r in reqs)
    assert eng.launches["meson_batches"] < eng.launches["meson_requests"]  # mesons were batched
    assert eng.launches["baryon_batches"] < eng.launches["baryon"]  # and baryons
    mixed = {r.rid: snapshot(torch, eng, r) for r in reqs}
    del eng
    for r in reqs:
        alone_req = serving.Request(r.rid, r.kind, r.dims, r.props, r.gammas, r.momenta)
        e1 = serve(torch, serving, [alone_req], cap_bytes=1 << 34, streams=1)
        alone = snapshot(torch, e1, alone_req)
        assert rel_err(mixed[r.rid], alone) <= 1e-12, (r.rid, r.kind)
    for r in reqs[:12]:
        assert rel_err(mixed[r.rid], oracle_result(torch, serving, r, cuda)) < 1e-12, (r.rid, r.kind)


def test_tight_pool_evicts_and_stays_correct(srv):
    """A cap of four small propagators: requests wait for memory, the LRU
    evicts between them, and every result still matches the oracle."""
    torch, serving, cuda = srv
    lat = [(4, 4, 4, 8)]
    reqs = serving.make_requests(2000.0, 0.01, seed=5, lattices=lat, catalogue=6, mix=(0.6, 0.4, 0.0),
                                 max_gammas=4, max_momenta=5)
    cap = 4 * 512 * 2304
    eng = serve(torch, serving, reqs, cap_bytes=cap, streams=3, max_inflight=4)
    assert eng.pool.evictions > 0 and eng.pool.peak <= cap
    for r in reqs:
        assert rel_err(r.result.cpu().numpy(), oracle_result(torch, serving, r, cuda)) < 1e-12, (r.rid, r.kind)


def test_realtime_stream_reports_latency(srv):
    torch, serving, cuda = srv
    reqs = serving.make_requests(400.0, 0.1, seed=2, lattices=LATTICES, catalogue=3, max_gammas=4, max_momenta=4)
    eng = serving.Engine(cap_bytes=1 << 34, streams=2, max_inflight=4)
    summary = eng.run(reqs)
    assert summary["completed"] == len(reqs)
    assert 0 < summary["p50_ms"] <= summary["p99_ms"]
    assert all(r.done >= r.arrival for r in reqs)
