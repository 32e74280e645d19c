"""The sequential source in FP32 site-scaled storage
(contract_baryon_sequential with CONTRACT_FP32_SCALED): all three buffers in
the site-scaled form, FP32 arithmetic, the target written back in that form.

* Every contraction type and target against the FP64 oracle run on the
  stored (decoded) values: per site, max |got - want| over the site's entries
  within 1e-6 of the site's largest |want|.
* A 64-timeslice exp(-0.8|t|) decay, where the sequential source spans ~44
  decades and plain complex64 products would underflow: the same bar.
* The FP64 form of the same entry point equals baryon_sequential_two_point
  bit for bit.
* The FP32 chain sequential source -> unit-Dirac sequential propagator ->
  three-point step, all in site-scaled FP32, within 1e-6 per timeslice of the
  FP64 oracle chain."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev(cuda):
    import torch
    from contract_amd import pycontract
    return torch, pycontract, cuda


def per_site_error(got, want):
    g = got.reshape(got.shape[0], -1)
    w = want.reshape(want.shape[0], -1)
    num = np.abs(g - w).max(axis=1)
    den = np.abs(w).max(axis=1)
    keep = den > 0
    return float((num[keep] / den[keep]).max())


def scaled_inputs(torch, pc, cuda, dims, alpha, seeds=(41, 42, 43)):
    ps = [oracle.synthetic_propagator(dims, s, alpha, t0=k) for k, s in enumerate(seeds)]
    sc = [pc.toScaledFP32(torch.from_numpy(p).to(cuda)) for p in ps]
    stored = [s.to_complex128().cpu().numpy() for s in sc]  # the values the FP32 form holds
    return sc, stored


@pytest.mark.parametrize("t", range(6))
@pytest.mark.parametrize("x", range(3))
def test_every_type_and_target(dev, t, x):
    torch, pc, cuda = dev
    dims = (4, 4, 2, 6)
    sc, stored = scaled_inputs(torch, pc, cuda, dims, 0.3)
    pc.baryonSequential(*sc, t, x, 5, 5, 8)
    got = sc[x].to_complex128().cpu().numpy()
    want = oracle.baryon_sequential_two_point(*stored, t, x, 5, 5, 8)[x]
    assert per_site_error(got, want) < 1e-6, (t, x)
    # the two inputs are untouched
    for k in range(3):
        if k != x:
            assert np.array_equal(sc[k].to_complex128().cpu().numpy(), stored[k])


@pytest.mark.parametrize("gammas", [(5, 5, 0), (10, 13, 15), (3, 5, 7)])
def test_other_gammas(dev, gammas):
    torch, pc, cuda = dev
    dims = (4, 2, 2, 4)
    sc, stored = scaled_inputs(torch, pc, cuda, dims, 0.2)
    pc.baryonSequential(*sc, 4, 1, *gammas)
    got = sc[1].to_complex128().cpu().numpy()
    want = oracle.baryon_sequential_two_point(*stored, 4, 1, *gammas)[1]
    assert per_site_error(got, want) < 1e-6


def test_steep_decay_keeps_every_site(dev):
    torch, pc, cuda = dev
    dims = (2, 2, 2, 64)
    sc, stored = scaled_inputs(torch, pc, cuda, dims, 0.8)
    pc.baryonSequential(*sc, 0, 2, 5, 5, 8)
    got = sc[2].to_complex128().cpu().numpy()
    want = oracle.baryon_sequential_two_point(*stored, 0, 2, 5, 5, 8)[2]
    mags = np.abs(want.reshape(want.shape[0], -1)).max(axis=1)
    assert mags.max() / mags[mags > 0].min() > 1e40  # the source spans > 40 decades
    assert per_site_error(got, want) < 1e-6


def test_fp64_form_equals_the_abi(dev):
    torch, pc, cuda = dev
    dims = (4, 4, 2, 4)
    ps = [oracle.synthetic_propagator(dims, s, 0.3) for s in (44, 45, 46)]
    a = [torch.from_numpy(p).to(cuda) for p in ps]
    b = [t.clone() for t in a]
    pc.baryonSequential(*a, 2, 0, 5, 5, 8)
    pc.baryonSequentialTwoPoint(*b, 2, 0, 5, 5, 8)
    torch.cuda.synchronize()
    assert torch.equal(a[0], b[0])


def test_invalid_arguments_are_reported(dev):
    torch, pc, cuda = dev
    dims = (4, 2, 2, 4)
    sc, _ = scaled_inputs(torch, pc, cuda, dims, 0.3)
    with pytest.raises(RuntimeError, match="sequential_type"):
        pc.baryonSequential(*sc, 0, 3, 5, 5, 8)
    with pytest.raises(RuntimeError, match="contract_type"):
        pc.baryonSequential(*sc, 6, 0, 5, 5, 8)
    with pytest.raises(TypeError):
        pc.baryonSequential(sc[0], sc[1], torch.zeros((64, 4, 4, 3, 3), dtype=torch.complex128, device=cuda),
                            0, 2, 5, 5, 8)


def test_fp32_sequential_chain_to_three_point(dev):
    torch, pc, cuda = dev
    dims = (4, 4, 4, 16)
    sc, stored = scaled_inputs(torch, pc, cuda, dims, 0.8)
    fwd = sc[2].clone()
    pc.baryonSequential(*sc, 0, 2, 5, 5, 8)
    sigma = pc.sequentialSource(sc[2])
    ins = [0, 4, 8, 15]
    moms = [(0, 0, 0), (1, 0, 0)]
    got = pc.threePointProject([(fwd, sigma, dims)], ins, moms)[0].cpu().numpy()
    seq = oracle.baryon_sequential_two_point(*stored, 0, 2, 5, 5, 8)[2]
    want = oracle.threepoint_project(stored[2], oracle.sequential_propagator_unit_dirac(seq), ins, dims, moms)
    num = np.abs(got - want).max(axis=(0, 1))
    den = np.abs(want).max(axis=(0, 1))
    assert (num / den).max() < 1e-6
