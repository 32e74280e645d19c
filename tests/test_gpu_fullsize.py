"""Full-size checks (the lattices of the benchmark configurations) through
size-independent properties; the oracle pins every path at small sizes.

* 24^3x64 proton (fused FP64 projection) = the drop-in per-site kernels'
  correlators summed over space with the phases computed independently in
  torch (FP64 matmul): 1e-12 per timeslice.
* 32^3x64: FP32 and FP32 site-scaled forms within 1e-6 per timeslice of the
  FP64 form, the site-scaled form bit-identical to the plain FP32 form.
* 32^3x64 meson grid request = per-site meson_two_point + torch projection
  for a few gamma pairs: 1e-12.
* Two 24^3x64 proton requests batched = the single calls, bit for bit."""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MOM27 = [(x, y, z) for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)]
MOM7 = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
PROTON = [(0, 1.0), (5, 1.0)]


@pytest.fixture(scope="module")
def dev(cuda):
    import torch
    from contract_amd import pycontract
    return torch, pycontract, cuda


def propagators(torch, dims, seeds, alpha, device):
    """iid N(0,1) entries times exp(-alpha |t|), lexicographic [V][4][4][3][3]."""
    Lx, Ly, Lz, Lt = dims
    prof = torch.exp(-alpha * torch.arange(Lt, dtype=torch.float64, device=device))
    out = []
    for s in seeds:
        g = torch.Generator(device=device)
        g.manual_seed(s)
        p = torch.view_as_complex(torch.randn((Lt, Lx * Ly * Lz, 144, 2), dtype=torch.float64, device=device,
                                              generator=g))
        p *= prof[:, None, None]
        out.append(p.reshape(-1, 4, 4, 3, 3).contiguous())
    return out


def torch_project(torch, corr, dims, moms):
    """[V] site correlator -> [nmom][Lt], phases exp(-2 pi i p.x / L) in FP64."""
    Lx, Ly, Lz, Lt = dims
    dev = corr.device
    x = torch.arange(Lx, device=dev, dtype=torch.float64)
    y = torch.arange(Ly, device=dev, dtype=torch.float64)
    z = torch.arange(Lz, device=dev, dtype=torch.float64)
    zz, yy, xx = torch.meshgrid(z, y, x, indexing="ij")
    P = torch.tensor(moms, dtype=torch.float64, device=dev)
    arg = (P[:, 0:1] * xx.reshape(1, -1) / Lx + P[:, 1:2] * yy.reshape(1, -1) / Ly
           + P[:, 2:3] * zz.reshape(1, -1) / Lz)
    ph = torch.polar(torch.ones_like(arg), -2 * math.pi * arg)  # [nmom][V3]
    c = corr.reshape(Lt, -1)
    return (ph @ c.transpose(0, 1).contiguous())  # [nmom][Lt]


def per_timeslice_error(got, want):
    num = np.abs(got - want).reshape(-1, got.shape[-1]).max(axis=0)
    den = np.abs(want).reshape(-1, want.shape[-1]).max(axis=0)
    return (num / np.maximum(den, 1e-300)).max()


def test_proton_24c64_fused_equals_site_kernels(dev):
    torch, pc, cuda = dev
    dims = (24, 24, 24, 64)
    u, d, u2 = propagators(torch, dims, (101, 102, 103), 0.5, cuda)
    gammas = [0, 5, 8, 15]
    fused = pc.baryonTwoPointProject(u, d, u2, PROTON, 5, 5, gammas, MOM27, dims).cpu().numpy()
    want = np.zeros_like(fused)
    for k, g in enumerate(gammas):
        corr = pc.baryonTwoPoint(u, d, u2, 0, 5, 5, g) + pc.baryonTwoPoint(u, d, u2, 5, 5, 5, g)
        want[k] = torch_project(torch, corr, dims, MOM27).cpu().numpy()
    assert per_timeslice_error(fused, want) < 1e-12


def test_fp32_forms_at_32c64(dev):
    torch, pc, cuda = dev
    dims = (32, 32, 32, 64)
    ps = propagators(torch, dims, (201, 202, 203), 0.5, cuda)
    gammas = list(range(16))
    ref = pc.baryonTwoPointProject(*ps, PROTON, 5, 5, gammas, MOM27, dims).cpu().numpy()
    sc = [pc.toScaledFP32(p) for p in ps]
    f32 = [p.to(torch.complex64) for p in ps]
    del ps
    got_s = pc.baryonTwoPointProject(*sc, PROTON, 5, 5, gammas, MOM27, dims).cpu().numpy()
    got_f = pc.baryonTwoPointProject(*f32, PROTON, 5, 5, gammas, MOM27, dims).cpu().numpy()
    assert per_timeslice_error(got_s, ref) < 1e-6
    assert np.array_equal(got_s, got_f)


def test_meson_grid_32c64_equals_site_kernel(dev):
    torch, pc, cuda = dev
    dims = (32, 32, 32, 64)
    pi, pj = propagators(torch, dims, (301, 302), 0.25, cuda)
    grid, = pc.mesonGridProject([(pi, pj, dims)], MOM7)
    grid = grid.cpu().numpy()
    V = pi.shape[0]
    corr = torch.empty(V, dtype=torch.complex128, device=cuda)
    for g1, g2 in [(0, 0), (5, 10), (15, 15), (8, 3)]:
        pc.meson_two_point(corr, pi, pj, V, g1, g2)
        want = torch_project(torch, corr, dims, MOM7).cpu().numpy()
        assert per_timeslice_error(grid[g1, g2], want) < 1e-12, (g1, g2)


def test_batch_of_two_24c64_equals_single_calls(dev):
    torch, pc, cuda = dev
    dims = (24, 24, 24, 64)
    a = propagators(torch, dims, (401, 402, 403), 0.5, cuda)
    b = propagators(torch, dims, (404, 405, 406), 0.5, cuda)
    outs = pc.baryonTwoPointProjectBatch([(*a, dims), (*b, dims)], PROTON, 5, 5, list(range(16)), MOM27)
    for props, got in zip((a, b), outs):
        alone = pc.baryonTwoPointProject(*props, PROTON, 5, 5, list(range(16)), MOM27, dims)
        assert torch.equal(got, alone)
