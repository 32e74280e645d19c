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
        t = eng.pool.get(r.result).tensor
        t = t.to_complex128() if hasattr(t, "to_complex128") else t
        return t.cpu().numpy().copy()
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
    """A cap of five small propagators: requests wait for memory, the LRU
    evicts between them, diquark outputs are copied out and released as
    their requests complete, and every result still matches the oracle."""
    torch, serving, cuda = srv
    lat = [(4, 4, 4, 8)]
    reqs = serving.make_requests(2000.0, 0.012, seed=5, lattices=lat, catalogue=6, mix=(0.5, 0.3, 0.2),
                                 max_gammas=4, max_momenta=5)
    assert {r.kind for r in reqs} == set(serving.KINDS)
    cap = 5 * 512 * 2304
    eng = serving.Engine(cap_bytes=cap, streams=3, max_inflight=4)
    got = {}

    def done(r):
        got[r.rid] = eng.result(r).cpu().numpy().copy()
        eng.release_result(r)

    eng.run(reqs, realtime=False, on_complete=done)
    assert eng.pool.evictions > 0 and eng.pool.peak <= cap and eng.pool.pinned_bytes() == 0
    for r in reqs:
        assert rel_err(got[r.rid], oracle_result(torch, serving, r, cuda)) < 1e-12, (r.rid, r.kind)


def test_unreleased_results_raise_instead_of_hanging(srv):
    """Diquark outputs kept pinned and never released fill a small pool: the
    loop reports it instead of waiting for memory that never frees."""
    torch, serving, cuda = srv
    k = [serving.PropKey((4, 4, 4, 8), i) for i in range(2)]
    reqs = [serving.Request(i, "diquark", (4, 4, 4, 8), (k[0], k[1]), [(5, 5)], []) for i in range(4)]
    eng = serving.Engine(cap_bytes=3 * 512 * 2304, streams=1, max_inflight=2)
    with pytest.raises(serving.PoolFull):
        eng.run(reqs, realtime=False)


def test_coalesced_requests_match_alone(srv):
    """Requests on the same propagators in one batch share one computation
    and each still gets its own gammas and momenta, equal to serving it
    alone."""
    torch, serving, cuda = srv
    dims = (4, 4, 4, 8)
    k = [serving.PropKey(dims, i) for i in range(2)]
    reqs = [serving.Request(0, "meson", dims, (k[0], k[1]), [(5, 5), (0, 8)], [(0, 0, 0), (1, 0, 0)]),
            serving.Request(1, "meson", dims, (k[0], k[1]), [(3, 3)], [(0, -1, 0)]),
            serving.Request(2, "baryon", dims, (k[0], k[1], k[0]), [0, 4], [(0, 0, 0)]),
            serving.Request(3, "baryon", dims, (k[0], k[1], k[0]), [8], [(0, 0, 1), (1, 1, 1)])]
    eng = serving.Engine(cap_bytes=1 << 30, streams=2, max_inflight=4)
    eng.run(reqs, realtime=False)
    assert eng.launches["meson_grids"] == 1 and eng.launches["baryon_computed"] == 1
    for r in reqs:
        alone = serving.Request(r.rid, r.kind, r.dims, r.props, r.gammas, r.momenta)
        e1 = serve(torch, serving, [alone], cap_bytes=1 << 30, streams=1)
        assert rel_err(r.result.cpu().numpy(), alone.result.cpu().numpy()) <= 1e-12
        assert rel_err(r.result.cpu().numpy(), oracle_result(torch, serving, r, cuda)) < 1e-12


def test_realtime_stream_reports_latency(srv):
    torch, serving, cuda = srv
    reqs = serving.make_requests(400.0, 0.1, seed=2, lattices=LATTICES, catalogue=3, max_gammas=4, max_momenta=4)
    eng = serving.Engine(cap_bytes=1 << 34, streams=2, max_inflight=4)
    summary = eng.run(reqs)
    assert summary["completed"] == len(reqs)
    assert 0 < summary["p50_ms"] <= summary["p99_ms"]
    assert all(r.done >= r.arrival for r in reqs)


def test_site_scaled_pool_matches_the_fp64_pool(srv):
    """The same request stream served from a pool of FP32 site-scaled
    propagators (packed from FP64 as they enter the pool; meson grids,
    baryon batches and diquarks all in that form) and from an FP64 pool:
    1e-6 relative per timeslice (projected results) and per propagator
    (diquarks), the FP32 bar.  The scaled pool's cap holds five of its
    entries, so scaled entries are evicted and their memory recycled."""
    torch, serving, cuda = srv
    lat = [(4, 4, 4, 8), (8, 4, 4, 4)]
    kw = dict(lattices=lat, catalogue=4, mix=(0.4, 0.3, 0.3), max_gammas=4, max_momenta=6)
    r64 = serving.make_requests(2000.0, 0.015, seed=9, **kw)
    r32 = serving.make_requests(2000.0, 0.015, seed=9, dtype="fp32_scaled", **kw)
    assert [r.kind for r in r64] == [r.kind for r in r32] and {r.kind for r in r64} == set(serving.KINDS)
    got = {}
    for name, reqs, cap in (("fp64", r64, 1 << 32), ("fp32_scaled", r32, 5 * 512 * 1153)):
        eng = serving.Engine(cap_bytes=cap, streams=3, max_inflight=4)
        out = got.setdefault(name, {})

        def done(r, eng=eng, out=out):
            t = eng.result(r)
            t = t.to_complex128() if hasattr(t, "to_complex128") else t
            out[r.rid] = t.cpu().numpy().copy()
            eng.release_result(r)

        eng.run(reqs, realtime=False, on_complete=done)
        if name == "fp32_scaled":
            assert eng.pool.evictions > 0 and eng.pool.peak <= cap
    for a, b in zip(r64, r32):
        assert rel_err(got["fp32_scaled"][b.rid], got["fp64"][a.rid]) < 1e-6, (a.rid, a.kind)
    # and the FP64 pool against the oracle, as in the other tests
    for r in r64[:6]:
        assert rel_err(got["fp64"][r.rid], oracle_result(torch, serving, r, cuda)) < 1e-12, (r.rid, r.kind)


def test_scaled_entry_holds_the_fp64_values(srv):
    """A propagator generated in the site-scaled form is the complex128
    propagator of the same key, packed (chunked generation included: the
    chunk size is forced small so the lattice spans several chunks)."""
    torch, serving, cuda = srv
    from contract_amd import pycontract as pc
    saved = serving.GEN_CHUNK_ENTRIES
    serving.GEN_CHUNK_ENTRIES = 3 * 64 * 144  # three 4x4x4 timeslices per chunk
    try:
        k64 = serving.PropKey((4, 4, 4, 8), 1, t0=3)
        k32 = serving.PropKey((4, 4, 4, 8), 1, t0=3, dtype="fp32_scaled")
        p64 = serving.device_propagator(torch, k64, cuda)
        p32 = serving.device_propagator(torch, k32, cuda)
        want = pc.toScaledFP32(p64)
        assert torch.equal(p32.data, want.data)
    finally:
        serving.GEN_CHUNK_ENTRIES = saved
