"""The fused baryon two-point with P_n the same buffer as P_i (one propagator
for a nucleon's two same-flavour quarks, (u, d, u)), where the library reads
and stages that propagator once per site.

Checked against the FP64 oracle, and against the same call with P_n an equal
copy in a buffer of its own (the three-propagator kernel): bit-identical in
FP64, FP32 and FP32 site-scaled storage.  The shared form reads the same
values in the same order and only moves the tails' LDS scratch; in FP32 it
reads P_n rescaled by P_i's power of two, which the site exponent undoes
exactly."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

MOM27 = [(x, y, z) for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)]
MOM7 = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
PROTON = [(0, 1.0), (5, 1.0)]  # IK_JL_NM + IM_JL_NK on (u, d, u)


def per_timeslice_error(got, want):
    num = np.abs(got - want).max(axis=(0, 1))
    den = np.abs(want).max(axis=(0, 1))
    return (num / np.maximum(den, 1e-300)).max()


@pytest.fixture(scope="module")
def dev(cuda):
    import torch
    from contract_amd import pycontract
    return torch, pycontract, cuda


def props(dev, dims, alpha=0.25, kind="fp64", seeds=(41, 42)):
    """(u, d, u_copy) on the device in the storage `kind`, and the complex128
    values the oracle sees (the FP32-stored values for the FP32 forms)."""
    torch, pc, cuda = dev
    host = [oracle.synthetic_propagator(dims, s, alpha, t0=1) for s in seeds]
    if kind == "fp32":
        host = [p.astype(np.complex64) for p in host]
    d = [torch.from_numpy(p).to(cuda) for p in host]
    if kind == "scaled":
        u, dq, u2 = pc.toScaledFP32(d[0]), pc.toScaledFP32(d[1]), pc.toScaledFP32(d[0].clone())
        return (u, dq, u2), [p.astype(np.complex128) for p in host]
    return (d[0], d[1], d[0].clone()), [p.astype(np.complex128) for p in host]


def shared_and_copy(dev, dims, terms, gammas, moms, order=0, kind="fp64", alpha=0.25, gij=5, gkl=5):
    torch, pc, _ = dev
    (u, d, u2), host = props(dev, dims, alpha, kind)
    shared = pc.baryonTwoPointProject(u, d, u, terms, gij, gkl, gammas, moms, dims, site_order=order)
    copy = pc.baryonTwoPointProject(u, d, u2, terms, gij, gkl, gammas, moms, dims, site_order=order)
    torch.cuda.synchronize()
    want = oracle.baryon_twopoint_project(host[0], host[1], host[0], terms, gij, gkl, gammas, dims, moms,
                                          order=order)
    return shared.cpu().numpy(), copy.cpu().numpy(), want


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("kind,tol", [("fp64", 1e-12), ("fp32", 1e-6), ("scaled", 1e-6)])
def test_proton_shared_u(dev, order, kind, tol):
    shared, copy, want = shared_and_copy(dev, (4, 4, 4, 6), PROTON, list(range(16)), MOM27, order, kind)
    assert np.array_equal(shared, copy)
    assert per_timeslice_error(shared, want) < tol


@pytest.mark.parametrize("t", range(6))
def test_every_type_and_a_projector(dev, t):
    """Each tail reads P_n in its own pattern; the type-0 tail also uses the
    LDS scratch, which the shared form moves to P_j's staging area."""
    half = 0.5 * (np.eye(4) - oracle.gamma(8))
    shared, copy, want = shared_and_copy(dev, (2, 4, 2, 4), [(t, 0.5 - 1.0j), ((t + 3) % 6, 0.25)], [7, half],
                                         MOM7, gkl=10)
    assert np.array_equal(shared, copy)
    assert per_timeslice_error(shared, want) < 1e-12


@pytest.mark.parametrize("nmom,slots", [(40, (17, 32)), (64, (33, 64))])
def test_two_and_more_momentum_tiles(dev, nmom, slots):
    """17-32 momentum slots (a momentum and its negative share one) take the
    shared form's second tile instance; more fall back to the
    three-propagator kernel, still with P_n = P_i."""
    rng = np.random.default_rng(nmom)
    moms = [(0, 0, 0)] + [tuple(int(v) for v in rng.integers(-3, 4, 3)) for _ in range(nmom - 1)]
    distinct = {min(m, tuple(-v for v in m)) for m in moms}
    assert slots[0] <= len(distinct) <= slots[1]
    shared, copy, want = shared_and_copy(dev, (6, 4, 2, 4), PROTON, [0, 15], moms)
    assert np.array_equal(shared, copy)
    assert per_timeslice_error(shared, want) < 1e-12


def test_fp32_steep_decay(dev):
    """exp(-0.8 |t - 1|) over 16 timeslices: the FP32 form's rescaling counts
    P_i's exponent twice when P_i serves as P_n."""
    shared, copy, want = shared_and_copy(dev, (2, 2, 2, 16), PROTON, [0, 5, 15], MOM7, kind="fp32", alpha=0.8)
    assert np.array_equal(shared, copy)
    assert per_timeslice_error(shared, want) < 1e-6


def test_batches_with_and_without_sharing(dev):
    """A launch takes the shared form only when every request shares; a mixed
    batch runs the three-propagator kernel.  Every request equals its single
    call either way."""
    torch, pc, cuda = dev
    lat = [(4, 4, 2, 4), (2, 4, 4, 6), (4, 2, 2, 8)]
    reqs_shared, reqs_mixed, singles = [], [], []
    for k, dims in enumerate(lat):
        (u, d, u2), _ = props(dev, dims, seeds=(50 + 2 * k, 51 + 2 * k))
        reqs_shared.append((u, d, u, dims))
        reqs_mixed.append((u, d, u if k != 1 else u2, dims))
        singles.append(pc.baryonTwoPointProject(u, d, u, PROTON, 5, 5, [0, 15], MOM27, dims))
    got_shared = pc.baryonTwoPointProjectBatch(reqs_shared, PROTON, 5, 5, [0, 15], MOM27)
    got_mixed = pc.baryonTwoPointProjectBatch(reqs_mixed, PROTON, 5, 5, [0, 15], MOM27)
    torch.cuda.synchronize()
    for s, a, b in zip(singles, got_shared, got_mixed):
        assert torch.equal(s, a) and torch.equal(s, b)


@pytest.mark.parametrize("kind,tol", [("fp64", 1e-12), ("scaled", 1e-6)])
@pytest.mark.parametrize("order", [0, 1])
def test_ragged_slices_shared(dev, kind, tol, order):
    """27 sites per timeslice (lexicographic 3^3x5) or per parity slice
    (even-odd 6x3x3x4): work items end part-way through a 4-site tile in the
    staged FP64 kernel and the FP32 forms, with P_n = P_i."""
    dims = (3, 3, 3, 5) if order == 0 else (6, 3, 3, 4)
    shared, copy, want = shared_and_copy(dev, dims, PROTON, [0, 8], MOM7, order, kind)
    assert np.array_equal(shared, copy)
    assert per_timeslice_error(shared, want) < tol
