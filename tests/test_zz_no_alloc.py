"""The fused baryon two-point for callers that cannot allocate
(contract_baryon_twopoint_project_noalloc), called through the C ABI:
results bit-identical to contract_baryon_twopoint_project with a library
workspace on two calls in a row (the second smaller, on the scratch the
first left behind), in a HIP graph replayed several times, and a call that
needs more scratch than the library holds is refused before any launch."""

import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

PROTON = [(0, 1.0), (5, 1.0)]
MOM27 = [(x, y, z) for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)]


@pytest.fixture(scope="module")
def dev(cuda):
    import torch
    from contract_amd import pycontract as pc
    pc.init(0)
    return torch, pc, cuda


def call_noalloc(pc, out, ps, insertions, moms, dims, stream=0):
    """contract_baryon_twopoint_project_noalloc through ctypes."""
    L = pc.lib()
    W = pc._projectors(insertions)
    mom = np.ascontiguousarray(moms, dtype=np.int32).reshape(-1, 3)
    types = np.ascontiguousarray([t for t, _ in PROTON], dtype=np.int32)
    w = np.ascontiguousarray([complex(x) for _, x in PROTON], dtype=np.complex128)
    lat = pc.Sublattice()
    for mu in range(4):
        lat.dims[mu] = int(dims[mu])
    vp = ctypes.c_void_p
    return L.contract_baryon_twopoint_project_noalloc(
        out.data_ptr(), ps[0].data_ptr(), ps[1].data_ptr(), ps[2].data_ptr(), 0, types.ctypes.data_as(vp),
        w.ctypes.data_as(vp), 2, 5, 5, W.ctypes.data_as(vp), W.shape[0], mom.ctypes.data_as(vp), mom.shape[0],
        ctypes.byref(lat), stream)


def case(torch, cuda, dims, seeds):
    return [torch.from_numpy(oracle.synthetic_propagator(dims, s, 0.4)).to(cuda) for s in seeds]


def test_two_calls_match_the_workspace_entry_point(dev):
    torch, pc, cuda = dev
    for dims, moms, gammas, seeds in (((16, 16, 16, 32), MOM27, list(range(16)), (11, 12, 13)),
                                      ((8, 4, 4, 8), MOM27[:7], [0, 8, 15], (14, 15, 16))):
        ps = case(torch, cuda, dims, seeds)
        ref = pc.baryonTwoPointProject(*ps, PROTON, 5, 5, gammas, moms, dims)
        out = torch.full_like(ref, float("nan"))
        assert call_noalloc(pc, out, ps, gammas, moms, dims) == 0, pc.last_error()
        torch.cuda.synchronize()
        assert torch.equal(out, ref), dims


def test_graph_replays_match_eager(dev):
    torch, pc, cuda = dev
    dims = (8, 8, 8, 8)
    ps = case(torch, cuda, dims, (21, 22, 21))
    gammas = [0, 4, 8, 15]
    eager = torch.empty((4, 27, 8), dtype=torch.complex128, device=cuda)
    assert call_noalloc(pc, eager, ps, gammas, MOM27, dims) == 0, pc.last_error()
    torch.cuda.synchronize()
    out = torch.empty_like(eager)
    side = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=side):
        rc = call_noalloc(pc, out, ps, gammas, MOM27, dims, stream=side.cuda_stream)
    assert rc == 0, pc.last_error()
    for _ in range(3):
        out.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager)
    want = oracle.baryon_twopoint_project(*[p.cpu().numpy() for p in ps], PROTON, 5, 5, gammas, dims, MOM27)
    assert np.abs(eager.cpu().numpy() - want).max() <= 1e-12 * np.abs(want).max()


def test_oversized_call_is_refused(dev):
    torch, pc, cuda = dev
    dims = (48, 48, 48, 96)
    moms = [(k, 0, 0) for k in range(1, 65)]  # 64 momenta, no +-k pairs
    reqs = [(None, None, None, dims)]
    need = pc.baryon_batch_workspace_bytes(reqs, len(moms))
    limit = pc.lib().contract_noalloc_scratch_bytes()
    assert need > limit  # the call below must stop at the size check, before any launch
    tiny = torch.zeros((4, 4, 4, 3, 3), dtype=torch.complex128, device=cuda)
    out = torch.zeros((1, 64, 96), dtype=torch.complex128, device=cuda)
    rc = call_noalloc(pc, out, [tiny, tiny, tiny], [8], moms, dims)
    assert rc == 3 and "scratch" in pc.last_error()
    assert torch.count_nonzero(out) == 0
