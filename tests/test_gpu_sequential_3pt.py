"""The sequential-source contraction on the GPU: building the sequential source
with the drop-in baryon_sequential_two_point and contracting it with a
sequential propagator per timeslice gives the fused two-point contraction with
that propagator in the removed slot (sum_x P_X . dC/dP_X = C).  This ties the
sequential entry point to the fused FP32/FP64 path used by the precision
study."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev(cuda):
    import torch
    from contract_amd import pycontract
    return torch, pycontract, cuda


@pytest.mark.parametrize("t", range(6))
@pytest.mark.parametrize("x", range(3))
def test_sequential_source_times_propagator_is_two_point(dev, t, x):
    torch, pc, cuda = dev
    dims = (4, 4, 2, 6)
    Lt = dims[3]
    ps = [oracle.synthetic_propagator(dims, s, 0.4, t0=2) for s in (61, 62, 63)]
    seqprop = oracle.synthetic_propagator(dims, 64, 0.4, t0=4)
    gmn = 8
    # sequential source in slot x, then contract with the sequential propagator
    d = [torch.from_numpy(p).to(cuda) for p in ps]
    pc.baryonSequentialTwoPoint(*d, t, x, 5, 5, gmn)
    src = d[x].reshape(Lt, -1)
    sp = torch.from_numpy(seqprop).to(cuda).reshape(Lt, -1)
    c3 = (src * sp).sum(dim=1).cpu().numpy()
    # fused two-point with the sequential propagator in slot x
    slots = [torch.from_numpy(p).to(cuda) for p in ps]
    slots[x] = torch.from_numpy(seqprop).to(cuda)
    fused = pc.baryonTwoPointProject(*slots, [(t, 1.0)], 5, 5, [gmn], [(0, 0, 0)], dims).cpu().numpy()[0, 0]
    per_t = np.abs(c3 - fused) / np.abs(fused)
    assert per_t.max() < 1e-12, per_t
