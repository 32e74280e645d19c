"""GPU parity of the drop-in ABI (include/contract.h) against the FP64 oracle.

Bar: FP64 results within 1e-12 of the oracle, relative to the largest
magnitude of the compared array (integer-exact layout, so any index slip
shows up as an O(1) error)."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def dev(cuda):
    import torch
    from contract_amd import pycontract
    pycontract.init(0)
    return torch, pycontract, cuda


@pytest.fixture(scope="module")
def small(golden):
    return golden["pi"], golden["pj"], golden["pn"]


@pytest.fixture(scope="module")
def medium():
    dims = (4, 4, 4, 8)
    return [oracle.synthetic_propagator(dims, s, 0.25, t0=3) for s in (21, 22, 23)]


def up(torch, cuda, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(cuda)


def test_meson_all_gamma_pairs(dev, small):
    torch, pc, cuda = dev
    pi, pj, _ = small
    di, dj = up(torch, cuda, pi), up(torch, cuda, pj)
    for gij in range(16):
        for gkl in range(16):
            got = pc.mesonTwoPoint(di, dj, gij, gkl).cpu().numpy()
            assert rel(got, oracle.meson_two_point(pi, pj, gij, gkl)) < TOL, (gij, gkl)


def test_meson_all_source_sink(dev, medium):
    torch, pc, cuda = dev
    pi, pj, _ = medium
    di, dj = up(torch, cuda, pi), up(torch, cuda, pj)
    for g in (0, 5, 8, 15):
        src = pc.mesonAllSourceTwoPoint(di, dj, g).cpu().numpy()
        snk = pc.mesonAllSinkTwoPoint(di, dj, g).cpu().numpy()
        assert rel(src, oracle.meson_all_source_two_point(pi, pj, g)) < TOL
        assert rel(snk, oracle.meson_all_sink_two_point(pi, pj, g)) < TOL


def test_diquark_all_gammas(dev, small):
    torch, pc, cuda = dev
    pi, pj, _ = small
    di, dj = up(torch, cuda, pi), up(torch, cuda, pj)
    for gij in range(16):
        for gkl in range(0, 16, 3):
            got = pc.baryonDiquark(di, dj, gij, gkl).cpu().numpy()
            assert rel(got, oracle.baryon_diquark(pi, pj, gij, gkl)) < TOL, (gij, gkl)


@pytest.mark.parametrize("t", range(6))
def test_baryon_every_type_and_insertion(dev, small, t):
    torch, pc, cuda = dev
    pi, pj, pn = small
    d = [up(torch, cuda, p) for p in small]
    for gij, gkl in [(5, 5), (13, 5), (10, 3), (0, 15)]:
        for gmn in range(16):
            got = pc.baryonTwoPoint(*d, t, gij, gkl, gmn).cpu().numpy()
            assert rel(got, oracle.baryon_two_point(pi, pj, pn, t, gij, gkl, gmn)) < TOL, (t, gij, gkl, gmn)


@pytest.mark.parametrize("t", range(6))
def test_baryon_general(dev, medium, t):
    torch, pc, cuda = dev
    pi, pj, pn = medium
    d = [up(torch, cuda, p) for p in medium]
    rng = np.random.default_rng(t)
    proj = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    half = 0.5 * (np.eye(4) - oracle.gamma(8))  # (1 - g4)/2, baryon_2pt.py:82
    for P in (proj, half):
        got = pc.baryonTwoPoint(*d, t, 5, 5, P).cpu().numpy()
        assert rel(got, oracle.baryon_general_two_point(pi, pj, pn, t, 5, 5, P)) < TOL


@pytest.mark.parametrize("t", range(6))
@pytest.mark.parametrize("s", range(3))
def test_sequential(dev, small, t, s):
    torch, pc, cuda = dev
    pi, pj, pn = small
    for gij, gkl, gmn in [(5, 5, 8), (13, 10, 3)]:
        d = [up(torch, cuda, p) for p in small]
        pc.baryonSequentialTwoPoint(*d, t, s, gij, gkl, gmn)
        want = oracle.baryon_sequential_two_point(pi, pj, pn, t, s, gij, gkl, gmn)
        for o in range(3):
            got = d[o].cpu().numpy()
            if o == s:
                assert rel(got, want[s]) < TOL, (t, s, gij, gkl, gmn)
            else:
                assert np.array_equal(got, small[o])


def test_larger_volume_and_empty(dev, medium):
    torch, pc, cuda = dev
    pi, pj, pn = medium
    d = [up(torch, cuda, p) for p in medium]
    got = pc.baryonTwoPoint(*d, 4, 5, 5, 8).cpu().numpy()
    assert rel(got, oracle.baryon_two_point(pi, pj, pn, 4, 5, 5, 8)) < TOL
    # volume 0 launches nothing
    out = torch.full((4,), 3.0 + 0j, dtype=torch.complex128, device=cuda)
    pc.baryon_two_point(out, d[0], d[1], d[2], 0, 0, 5, 5, 8)
    pc.meson_two_point(out, d[0], d[1], 0, 5, 5)
    torch.cuda.synchronize()
    assert (out.cpu().numpy() == 3.0).all()


@pytest.fixture(scope="module")
def multi_tile():
    """16^4 sites: 16384 four-site tiles, several per wave on any MI355X
    launch (at most 8 waves per CU), so the kernels' steady-state pipelines
    (next-tile loads in flight across steps) are exercised, not only the
    first tile of a run."""
    dims = (16, 16, 16, 16)
    return [oracle.synthetic_propagator(dims, s, 0.25, t0=5) for s in (31, 32, 33)]


def test_multi_tile_runs(dev, multi_tile):
    torch, pc, cuda = dev
    pi, pj, pn = multi_tile
    d = [up(torch, cuda, p) for p in multi_tile]
    for t in range(6):
        got = pc.baryonTwoPoint(*d, t, 5, 5, 8).cpu().numpy()
        assert rel(got, oracle.baryon_two_point(pi, pj, pn, t, 5, 5, 8)) < TOL, t
    got = pc.baryonTwoPoint(*d, 2, 13, 10, 3).cpu().numpy()  # runtime Gij form
    assert rel(got, oracle.baryon_two_point(pi, pj, pn, 2, 13, 10, 3)) < TOL
    half = 0.5 * (np.eye(4) - oracle.gamma(8))
    got = pc.baryonTwoPoint(*d, 0, 5, 5, half).cpu().numpy()
    assert rel(got, oracle.baryon_general_two_point(pi, pj, pn, 0, 5, 5, half)) < TOL
    got = pc.mesonTwoPoint(d[0], d[1], 5, 8).cpu().numpy()
    assert rel(got, oracle.meson_two_point(pi, pj, 5, 8)) < TOL
    src = pc.mesonAllSourceTwoPoint(d[0], d[1], 5).cpu().numpy()
    snk = pc.mesonAllSinkTwoPoint(d[0], d[1], 8).cpu().numpy()
    assert rel(src, oracle.meson_all_source_two_point(pi, pj, 5)) < TOL
    assert rel(snk, oracle.meson_all_sink_two_point(pi, pj, 8)) < TOL
    got = pc.baryonDiquark(d[0], d[1], 5, 5).cpu().numpy()
    assert rel(got, oracle.baryon_diquark(pi, pj, 5, 5)) < TOL
    pc.baryonSequentialTwoPoint(*d, 0, 2, 5, 5, 8)
    want = oracle.baryon_sequential_two_point(pi, pj, pn, 0, 2, 5, 5, 8)
    assert rel(d[2].cpu().numpy(), want[2]) < TOL


def test_extreme_magnitudes(dev):
    """Steep exp(-0.8|t|) decay: ~20 decades across the lattice, still FP64-exact."""
    torch, pc, cuda = dev
    dims = (2, 2, 2, 16)
    ps = [oracle.synthetic_propagator(dims, s, 0.8) for s in (4, 5, 6)]
    d = [up(torch, cuda, p) for p in ps]
    got = pc.baryonTwoPoint(*d, 1, 5, 5, 8).cpu().numpy()
    want = oracle.baryon_two_point(*ps, 1, 5, 5, 8)
    per_t = np.abs(got - want).reshape(16, -1).max(1) / np.abs(want).reshape(16, -1).max(1)
    assert per_t.max() < TOL
