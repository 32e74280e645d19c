"""GPU parity of the batched fused baryon two-point
(contract_baryon_twopoint_project_batch): several propagator triples --
source positions, flavour combinations, lattices of different size and site
order, sub-lattices of a larger grid -- in one launch.

Each request's output must match the FP64 oracle to 1e-12 per timeslice
(relative, max over insertions and momenta) and be bit-identical to the same
request run alone: the batch only concatenates work items, so every partial
and every reduction order is the request's own."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

MOM27 = [(x, y, z) for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)]
PROTON = [(0, 1.0), (5, 1.0)]


def per_timeslice_error(got, want):
    num = np.abs(got - want).max(axis=(0, 1))
    den = np.abs(want).max(axis=(0, 1))
    return (num / np.maximum(den, 1e-300)).max()


@pytest.fixture(scope="module")
def dev(cuda):
    import torch
    from contract_amd import pycontract
    return torch, pycontract, cuda


# (dims, site_order, global spatial extents or None, origin or None)
UNEVEN = [
    ((4, 4, 4, 8), 0, None, None),
    ((6, 4, 2, 6), 1, None, None),
    ((3, 3, 3, 5), 0, None, None),
    ((4, 4, 2, 4), 1, (10, 6, 4), (2, 4, 0)),
    ((8, 4, 4, 4), 0, None, None),
    ((2, 2, 4, 3), 0, (6, 6, 4), (3, 1, 0)),
]


def make_batch(torch, cuda, cases, seed0, fp32=False):
    host, reqs = [], []
    for n, (dims, order, G, origin) in enumerate(cases):
        ps = [oracle.synthetic_propagator(dims, seed0 + 3 * n + k, 0.25, t0=n % dims[3]) for k in range(3)]
        if fp32:
            ps = [p.astype(np.complex64) for p in ps]
        host.append(ps)
        reqs.append((*[torch.from_numpy(p).to(cuda) for p in ps], dims, order, G, origin))
    return host, reqs


def test_uneven_batch_matches_oracle_and_single_calls(dev):
    torch, pc, cuda = dev
    gammas = [0, 5, 8, 15]
    host, reqs = make_batch(torch, cuda, UNEVEN, 100)
    outs = [o.cpu().numpy() for o in pc.baryonTwoPointProjectBatch(reqs, PROTON, 5, 5, gammas, MOM27)]
    for (Pi, Pj, Pn, dims, order, G, origin), ps, got in zip(reqs, host, outs):
        assert got.shape == (4, 27, dims[3])
        alone = pc.baryonTwoPointProject(Pi, Pj, Pn, PROTON, 5, 5, gammas, MOM27, dims, site_order=order,
                                         global_dims=G, origin=origin).cpu().numpy()
        assert np.array_equal(got, alone), dims
        want = oracle.baryon_twopoint_project(*ps, PROTON, 5, 5, gammas, dims, MOM27, order=order, global_dims=G,
                                              origin=origin)
        assert per_timeslice_error(got, want) < 1e-12, dims


def test_more_requests_than_one_launch(dev):
    """18 requests: two launches (16 + 2), runtime diquark gamma, general projector."""
    torch, pc, cuda = dev
    cases = [((2, 4, 2, 4), n % 2, None, None) for n in range(18)]
    host, reqs = make_batch(torch, cuda, cases, 500)
    half = 0.5 * (np.eye(4) - oracle.gamma(8))
    terms = [(2, 1.0), (4, -0.5j)]
    moms = MOM27[:7]
    outs = pc.baryonTwoPointProjectBatch(reqs, terms, 13, 10, [7, half], moms)
    for (Pi, Pj, Pn, dims, order, _, _), ps, got in zip(reqs, host, outs):
        want = oracle.baryon_twopoint_project(*ps, terms, 13, 10, [7, half], dims, moms, order=order)
        assert per_timeslice_error(got.cpu().numpy(), want) < 1e-12


def test_fp32_batch_equals_fp32_single_calls(dev):
    torch, pc, cuda = dev
    host, reqs = make_batch(torch, cuda, UNEVEN[:4], 900, fp32=True)
    outs = pc.baryonTwoPointProjectBatch(reqs, PROTON, 5, 5, list(range(16)), MOM27)
    for (Pi, Pj, Pn, dims, order, G, origin), ps, got in zip(reqs, host, outs):
        alone = pc.baryonTwoPointProject(Pi, Pj, Pn, PROTON, 5, 5, list(range(16)), MOM27, dims, site_order=order,
                                         global_dims=G, origin=origin)
        assert torch.equal(got, alone)
        want = oracle.baryon_twopoint_project(*[p.astype(np.complex128) for p in ps], PROTON, 5, 5, list(range(16)),
                                              dims, MOM27, order=order, global_dims=G, origin=origin)
        assert per_timeslice_error(got.cpu().numpy(), want) < 1e-6


def test_caller_workspace(dev):
    torch, pc, cuda = dev
    _, reqs = make_batch(torch, cuda, UNEVEN[:3], 40)
    nbytes = pc.baryon_batch_workspace_bytes(reqs, len(MOM27))
    ws = torch.empty(nbytes, dtype=torch.uint8, device=cuda)
    a = pc.baryonTwoPointProjectBatch(reqs, PROTON, 5, 5, [0, 8], MOM27, workspace=ws)
    b = pc.baryonTwoPointProjectBatch(reqs, PROTON, 5, 5, [0, 8], MOM27)
    assert all(torch.equal(x, y) for x, y in zip(a, b))
    with pytest.raises(RuntimeError, match="workspace"):
        pc.baryonTwoPointProjectBatch(reqs, PROTON, 5, 5, [0, 8], MOM27, workspace=ws[:4096])


@pytest.mark.parametrize("scaled", [False, True])
def test_graph_replay(dev, scaled):
    """The batched fused launch with a caller workspace allocates nothing, so
    it can be captured in a HIP graph; every replay (with eager work and
    synchronisation in between) matches the eager result bit for bit."""
    torch, pc, cuda = dev
    _, reqs = make_batch(torch, cuda, UNEVEN[:4], 70)
    if scaled:
        reqs = [(*[pc.toScaledFP32(p) for p in r[:3]], *r[3:]) for r in reqs]
    nbytes = pc.baryon_batch_workspace_bytes(reqs, len(MOM27))
    ws = torch.empty(nbytes, dtype=torch.uint8, device=cuda)
    outs = [torch.empty((16, 27, r[3][3]), dtype=torch.complex128, device=cuda) for r in reqs]
    eager = [o.cpu().numpy() for o in pc.baryonTwoPointProjectBatch(reqs, PROTON, 5, 5, list(range(16)), MOM27,
                                                                    workspace=ws)]
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        pc.baryonTwoPointProjectBatch(reqs, PROTON, 5, 5, list(range(16)), MOM27, outs=outs, workspace=ws,
                                      stream=s.cuda_stream)
    for rep in range(3):
        for o in outs:
            o.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for o, e in zip(outs, eager):
            assert np.array_equal(o.cpu().numpy(), e), rep
