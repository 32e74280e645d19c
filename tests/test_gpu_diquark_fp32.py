"""The diquark in FP32 site-scaled storage (contract_baryon_diquark with
CONTRACT_FP32_SCALED): both propagators and the output in the site-scaled
form, FP32 arithmetic.

* Several (Gij, Gkl) against the FP64 oracle run on the stored (decoded)
  values: per site, max |got - want| within 1e-6 of the site's largest
  |want|.
* A 64-timeslice exp(-0.8|t|) decay, where the diquark spans > 40 decades
  and plain complex64 products would underflow: the same bar.
* The FP64 form of the same entry point equals baryon_diquark bit for bit.
* Invalid arguments are reported, not launched."""

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


def scaled_pair(torch, pc, cuda, dims, alpha, seeds=(51, 52)):
    ps = [oracle.synthetic_propagator(dims, s, alpha, t0=k) for k, s in enumerate(seeds)]
    sc = [pc.toScaledFP32(torch.from_numpy(p).to(cuda)) for p in ps]
    stored = [s.to_complex128().cpu().numpy() for s in sc]  # the values the FP32 form holds
    return sc, stored


def empty_scaled(torch, pc, cuda, V):
    return pc.ScaledPropagator(torch.empty(pc.scaled_fp32_bytes(V), dtype=torch.uint8, device=cuda), V)


@pytest.mark.parametrize("gammas", [(5, 5), (10, 13), (0, 15), (3, 7)])
def test_against_oracle(dev, gammas):
    torch, pc, cuda = dev
    dims = (4, 4, 2, 6)
    V = int(np.prod(dims))
    sc, stored = scaled_pair(torch, pc, cuda, dims, 0.3)
    out = empty_scaled(torch, pc, cuda, V)
    pc.baryonDiquarkInto(out, *sc, *gammas)
    got = out.to_complex128().cpu().numpy()
    want = oracle.baryon_diquark(*stored, *gammas)
    assert per_site_error(got, want) < 1e-6, gammas
    for k in range(2):  # the inputs are untouched
        assert np.array_equal(sc[k].to_complex128().cpu().numpy(), stored[k])


def test_steep_decay_keeps_every_site(dev):
    torch, pc, cuda = dev
    dims = (2, 2, 2, 64)
    V = int(np.prod(dims))
    sc, stored = scaled_pair(torch, pc, cuda, dims, 0.8)
    out = empty_scaled(torch, pc, cuda, V)
    pc.baryonDiquarkInto(out, *sc, 5, 5)
    got = out.to_complex128().cpu().numpy()
    want = oracle.baryon_diquark(*stored, 5, 5)
    mags = np.abs(want.reshape(V, -1)).max(axis=1)
    assert mags.max() / mags[mags > 0].min() > 1e40
    assert per_site_error(got, want) < 1e-6


def test_ragged_volume(dev):
    # a volume that is not a multiple of the 4-site tile
    torch, pc, cuda = dev
    V = 4 * 6 + 2
    ps = [oracle.synthetic_propagator((V, 1, 1, 1), s, 0.0) for s in (61, 62)]
    sc = [pc.toScaledFP32(torch.from_numpy(p).to(cuda)) for p in ps]
    stored = [s.to_complex128().cpu().numpy() for s in sc]
    out = empty_scaled(torch, pc, cuda, V)
    pc.baryonDiquarkInto(out, *sc, 5, 5)
    assert per_site_error(out.to_complex128().cpu().numpy(), oracle.baryon_diquark(*stored, 5, 5)) < 1e-6


def test_fp64_form_equals_the_abi(dev):
    torch, pc, cuda = dev
    dims = (4, 4, 2, 4)
    V = int(np.prod(dims))
    pi, pj = (torch.from_numpy(oracle.synthetic_propagator(dims, s, 0.3)).to(cuda) for s in (63, 64))
    a = torch.empty_like(pi)
    b = torch.empty_like(pi)
    pc.baryonDiquarkInto(a, pi, pj, 5, 10)
    pc.baryon_diquark(b, pi, pj, V, 5, 10)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


def test_invalid_arguments_are_reported(dev):
    torch, pc, cuda = dev
    dims = (4, 2, 2, 4)
    V = int(np.prod(dims))
    sc, _ = scaled_pair(torch, pc, cuda, dims, 0.3)
    out = empty_scaled(torch, pc, cuda, V)
    with pytest.raises(RuntimeError, match="gamma_ij"):
        pc.baryonDiquarkInto(out, *sc, 16, 5)
    with pytest.raises(TypeError):
        pc.baryonDiquarkInto(torch.zeros((V, 4, 4, 3, 3), dtype=torch.complex128, device=cuda), *sc, 5, 5)
    small = empty_scaled(torch, pc, cuda, V - 4)
    with pytest.raises(ValueError):
        pc.baryonDiquarkInto(small, *sc, 5, 5)
