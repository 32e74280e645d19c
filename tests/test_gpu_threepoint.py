"""GPU parity of the three-point step after the sequential source
(contract_threepoint_project, include/contract_ext.h).

* against the FP64 oracle's closed form Tr[g5 Gamma F Sigma^dagger], projected
  (lexicographic and even-odd order, gamma indices and a dense insertion,
  +-k momenta, more than one momentum tile, sub-lattices);
* the unit-Dirac-operator chain, which ties the step to the drop-in
  sequential source: with D = 1 the sequential propagator is the solver
  source sigma = g5 conj(seq_X) itself, and the three-point function of Gamma
  equals the two-point function with P_X replaced by Gamma P_X, for all six
  contraction types and all three targets (the library's sequential source,
  three-point step and fused two-point against each other);
* FP32 and site-scaled FP32 storage against the FP64 oracle on the stored
  values, including steep decays;
* batches bit-identical to single calls.

Bars: 1e-12 relative per timeslice (FP64), 1e-6 (FP32 storage); per-timeslice
relative error = max over insertions and momenta of |got - want| over the
largest |want| of the timeslice."""

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

MOM7 = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
DENSE = [[0.5, 1j, 0, 0], [0, -1, 0, 2], [0.25, 0, 1, 0], [0, 0, -1j, 3]]


def per_timeslice_error(got, want):
    num = np.abs(got - want).max(axis=(0, 1))
    den = np.abs(want).max(axis=(0, 1))
    return float((num / np.maximum(den, 1e-300)).max())


@pytest.fixture(scope="module")
def dev(cuda):
    import torch
    from contract_amd import pycontract
    return torch, pycontract, cuda


def gpu(torch, cuda, *arrays):
    return [torch.from_numpy(np.ascontiguousarray(a)).to(cuda) for a in arrays]


@pytest.mark.parametrize("order", [0, 1])
def test_against_oracle(dev, order):
    torch, pc, cuda = dev
    dims = (4, 6, 2, 6)
    f = oracle.synthetic_propagator(dims, 71, 0.3, t0=1)
    s = oracle.synthetic_propagator(dims, 72, 0.3, t0=4)
    ins = [0, 1, 7, 8, 15, DENSE]
    F, S = gpu(torch, cuda, f, s)
    got = pc.threePointProject([(F, S, dims, order)], ins, MOM7)[0].cpu().numpy()
    want = oracle.threepoint_project(f, s, ins, dims, MOM7, order=order)
    assert got.shape == (len(ins), 7, dims[3])
    assert per_timeslice_error(got, want) < 1e-12


@pytest.mark.parametrize("nmom", [1, 20, 40, 64])
def test_momentum_tiles(dev, nmom):
    torch, pc, cuda = dev
    dims = (6, 4, 4, 4)
    rng = np.random.default_rng(nmom)
    moms = [(0, 0, 0)] + [tuple(int(v) for v in rng.integers(-3, 4, 3)) for _ in range(nmom - 1)]
    f = oracle.synthetic_propagator(dims, 73, 0.2)
    s = oracle.synthetic_propagator(dims, 74, 0.2)
    F, S = gpu(torch, cuda, f, s)
    got = pc.threePointProject([(F, S, dims)], [3, 12], moms)[0].cpu().numpy()
    want = oracle.threepoint_project(f, s, [3, 12], dims, moms)
    assert per_timeslice_error(got, want) < 1e-12


def test_equals_meson_with_g5_source(dev):
    """Per gamma index the step is meson_two_point(F, Sigma; Gamma, 15): the
    drop-in meson kernel, projected by the oracle, agrees."""
    torch, pc, cuda = dev
    dims = (4, 4, 4, 4)
    f = oracle.synthetic_propagator(dims, 75, 0.2)
    s = oracle.synthetic_propagator(dims, 76, 0.2)
    F, S = gpu(torch, cuda, f, s)
    got = pc.threePointProject([(F, S, dims)], list(range(16)), MOM7[:3])[0].cpu().numpy()
    site = np.stack([pc.mesonTwoPoint(F, S, g, 15).cpu().numpy() for g in range(16)], axis=1)
    want = oracle.project(site, dims, MOM7[:3])
    assert per_timeslice_error(got, want) < 1e-12


@pytest.mark.parametrize("t", range(6))
@pytest.mark.parametrize("x", range(3))
def test_unit_dirac_chain(dev, t, x):
    """seq_X from the library's baryon_sequential_two_point; Sigma = sigma
    (D = 1); the three-point of Gamma equals the fused two-point with P_X
    replaced by Gamma P_X."""
    torch, pc, cuda = dev
    dims = (4, 4, 2, 6)
    ps = [oracle.synthetic_propagator(dims, s, 0.4, t0=2) for s in (81, 82, 83)]
    d = gpu(torch, cuda, *ps)
    fwd = d[x].clone()
    pc.baryonSequentialTwoPoint(*d, t, x, 5, 5, 8)
    sigma = pc.sequentialSource(d[x])
    ins = [0, 2, 8, 11, 15]
    moms = MOM7[:3]
    got = pc.threePointProject([(fwd, sigma, dims)], ins, moms)[0].cpu().numpy()
    for w, g in enumerate(ins):
        G = torch.from_numpy(oracle.gamma(g)).to(cuda)
        slots = [torch.from_numpy(p).to(cuda) for p in ps]
        slots[x] = torch.einsum("ab,xbscd->xascd", G, slots[x]).contiguous()
        want = pc.baryonTwoPointProject(*slots, [(t, 1.0)], 5, 5, [8], moms, dims).cpu().numpy()
        assert per_timeslice_error(got[w:w + 1], want) < 1e-12, (t, x, g)


def test_sublattice_halves_add_up(dev):
    torch, pc, cuda = dev
    dims = (8, 4, 4, 4)
    f = oracle.synthetic_propagator(dims, 77, 0.2)
    s = oracle.synthetic_propagator(dims, 78, 0.2)
    moms = [(1, 0, 0), (-1, 1, 0), (0, 0, 0)]
    whole = oracle.threepoint_project(f, s, [8, 15], dims, moms)
    fr = f.reshape(4, 4, 4, 8, -1)
    sr = s.reshape(4, 4, 4, 8, -1)
    parts = []
    for x0, nx in ((0, 2), (2, 6)):
        sub = (nx, 4, 4, 4)
        fs = np.ascontiguousarray(fr[:, :, :, x0:x0 + nx]).reshape(-1, 4, 4, 3, 3)
        ss = np.ascontiguousarray(sr[:, :, :, x0:x0 + nx]).reshape(-1, 4, 4, 3, 3)
        F, S = gpu(torch, cuda, fs, ss)
        parts.append(pc.threePointProject([(F, S, sub, 0, (8, 4, 4, 4), (x0, 0, 0, 0))], [8, 15], moms)[0]
                     .cpu().numpy())
    assert per_timeslice_error(parts[0] + parts[1], whole) < 1e-12


@pytest.mark.parametrize("form", ["fp32", "fp32_scaled"])
@pytest.mark.parametrize("alpha", [0.3, 0.8])
def test_fp32_storage(dev, form, alpha):
    torch, pc, cuda = dev
    dims = (4, 4, 4, 32)
    f = oracle.synthetic_propagator(dims, 79, alpha).astype(np.complex64)
    s = oracle.synthetic_propagator(dims, 80, alpha, t0=31).astype(np.complex64)
    F, S = gpu(torch, cuda, f, s)
    if form == "fp32_scaled":
        F, S = pc.toScaledFP32(F.to(torch.complex128)), pc.toScaledFP32(S.to(torch.complex128))
    ins = [0, 4, 8, 15]
    got = pc.threePointProject([(F, S, dims)], ins, MOM7)[0].cpu().numpy()
    want = oracle.threepoint_project(f.astype(np.complex128), s.astype(np.complex128), ins, dims, MOM7)
    assert per_timeslice_error(got, want) < 1e-6


def test_batch_is_bit_identical_to_singles(dev):
    torch, pc, cuda = dev
    lats = [(4, 4, 4, 8), (6, 4, 2, 6), (8, 4, 4, 4), (4, 4, 4, 8)]
    reqs = []
    for k, dims in enumerate(lats):
        f = oracle.synthetic_propagator(dims, 90 + 2 * k, 0.3)
        s = oracle.synthetic_propagator(dims, 91 + 2 * k, 0.3)
        reqs.append((*gpu(torch, cuda, f, s), dims, 1 if dims[0] % 2 == 0 and k == 3 else 0))
    ins = [1, 8, DENSE]
    got = [o.cpu() for o in pc.threePointProject(reqs, ins, MOM7)]
    for r, g in zip(reqs, got):
        alone = pc.threePointProject([r], ins, MOM7)[0].cpu()
        assert torch.equal(alone, g)
        want = oracle.threepoint_project(r[0].cpu().numpy(), r[1].cpu().numpy(), ins, r[2], MOM7, order=r[3])
        assert per_timeslice_error(g.numpy(), want) < 1e-12


def test_workspace_and_graph_replay(dev):
    """With a caller workspace the step allocates nothing and replays from a
    HIP graph, every replay equal to the eager call."""
    torch, pc, cuda = dev
    dims = (8, 8, 8, 8)
    f = oracle.synthetic_propagator(dims, 95, 0.3)
    s = oracle.synthetic_propagator(dims, 96, 0.3)
    F, S = gpu(torch, cuda, f, s)
    req = [(F, S, dims)]
    ws = torch.empty(pc.threepoint_workspace_bytes(req, 7), dtype=torch.uint8, device=cuda)
    out = [torch.empty((16, 7, 8), dtype=torch.complex128, device=cuda)]
    pc.threePointProject(req, list(range(16)), MOM7, outs=out, workspace=ws)
    eager = out[0].clone()
    side = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=side):
        pc.threePointProject(req, list(range(16)), MOM7, outs=out, workspace=ws, stream=side.cuda_stream)
    for _ in range(3):
        out[0].zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out[0], eager)
    want = oracle.threepoint_project(f, s, list(range(16)), dims, MOM7)
    assert per_timeslice_error(eager.cpu().numpy(), want) < 1e-12
