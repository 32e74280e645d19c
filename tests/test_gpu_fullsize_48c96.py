"""Checks on the largest lattice of the serving mix, 48^3x96 (10,616,832 sites,
24.5 GB per FP64 propagator: every buffer spans more than 2^32 bytes), through
properties that hold at any size; the oracle pins each path at small sizes.

* fused proton two-point (three buffers, FP64) = the drop-in per-site
  baryon_two_point of both terms, projected with phases computed
  independently in torch (FP64 matmul): 1e-12 per timeslice;
* the FP32 site-scaled form of the same step (its exponents sit 12.2 GB into
  each buffer) within 1e-6 per timeslice of FP64;
* one meson-grid request = per-site meson_two_point + torch projection for a
  few gamma pairs: 1e-12;
* the three-point step = per-site meson_two_point(F, Sigma; Gamma, 15) +
  torch projection: 1e-12;
* baryon_diquark and baryon_sequential_two_point on the whole lattice, read
  back at sites on both sides of the 2^31- and 2^32-byte offsets and at the
  end of the buffer, = the oracle on those sites: 1e-12."""

import math

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

DIMS = (48, 48, 48, 96)
V = 48 ** 3 * 96
MOM27 = [(x, y, z) for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)]
MOM7 = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
PROTON = [(0, 1.0), (5, 1.0)]
# sites whose 2304-byte blocks straddle the 2^31 and 2^32 byte offsets, the
# first and the last sites
EDGE_SITES = np.array([0, 1, 2, 3, 932064, 932065, 932066, 932067, 932068, 932069,
                       1864132, 1864133, 1864134, 1864135, 1864136, 1864137, V - 6, V - 5, V - 4, V - 3, V - 2, V - 1])


@pytest.fixture(scope="module")
def big(cuda):
    import torch
    from contract_amd import pycontract as pc
    prof = torch.exp(-0.3 * torch.arange(DIMS[3], dtype=torch.float64, device=cuda))
    ps = []
    for s in (4801, 4802, 4803):
        g = torch.Generator(device=cuda)
        g.manual_seed(s)
        p = torch.view_as_complex(torch.randn((DIMS[3], V // DIMS[3], 144, 2), dtype=torch.float64, device=cuda,
                                              generator=g))
        p *= prof[:, None, None]
        ps.append(p.reshape(V, 4, 4, 3, 3))
    yield torch, pc, cuda, ps
    del ps
    torch.cuda.empty_cache()


def torch_project(torch, corr, moms):
    Lx, Ly, Lz, Lt = DIMS
    dev = corr.device
    zz, yy, xx = torch.meshgrid(*(torch.arange(n, device=dev, dtype=torch.float64) for n in (Lz, Ly, Lx)),
                                indexing="ij")
    P = torch.tensor(moms, dtype=torch.float64, device=dev)
    arg = P[:, 0:1] * xx.reshape(1, -1) / Lx + P[:, 1:2] * yy.reshape(1, -1) / Ly + P[:, 2:3] * zz.reshape(1, -1) / Lz
    ph = torch.polar(torch.ones_like(arg), -2 * math.pi * arg)
    return ph @ corr.reshape(Lt, -1).transpose(0, 1).contiguous()


def per_timeslice_error(got, want):
    num = np.abs(got - want).reshape(-1, got.shape[-1]).max(axis=0)
    den = np.abs(want).reshape(-1, want.shape[-1]).max(axis=0)
    return float((num / np.maximum(den, 1e-300)).max())


def test_proton_fused_equals_site_kernels(big):
    torch, pc, cuda, (u, d, u2) = big
    gammas = [0, 8]
    fused = pc.baryonTwoPointProject(u, d, u2, PROTON, 5, 5, gammas, MOM27, DIMS).cpu().numpy()
    want = np.zeros_like(fused)
    for k, g in enumerate(gammas):
        corr = pc.baryonTwoPoint(u, d, u2, 0, 5, 5, g) + pc.baryonTwoPoint(u, d, u2, 5, 5, 5, g)
        want[k] = torch_project(torch, corr, MOM27).cpu().numpy()
    assert per_timeslice_error(fused, want) < 1e-12


def test_scaled_fp32_form(big):
    torch, pc, cuda, (u, d, u2) = big
    gammas = [0, 8, 15]
    ref = pc.baryonTwoPointProject(u, d, u2, PROTON, 5, 5, gammas, MOM7, DIMS).cpu().numpy()
    sc = [pc.toScaledFP32(p) for p in (u, d, u2)]
    got = pc.baryonTwoPointProject(*sc, PROTON, 5, 5, gammas, MOM7, DIMS).cpu().numpy()
    del sc
    torch.cuda.empty_cache()
    assert per_timeslice_error(got, ref) < 1e-6


def test_meson_grid_equals_site_kernel(big):
    torch, pc, cuda, (u, d, _) = big
    grid, = pc.mesonGridProject([(u, d, DIMS)], MOM7)
    grid = grid.cpu().numpy()
    corr = torch.empty(V, dtype=torch.complex128, device=cuda)
    for g1, g2 in [(0, 0), (5, 10), (15, 15)]:
        pc.meson_two_point(corr, u, d, V, g1, g2)
        assert per_timeslice_error(grid[g1, g2], torch_project(torch, corr, MOM7).cpu().numpy()) < 1e-12, (g1, g2)


def test_threepoint_equals_meson_with_g5_source(big):
    torch, pc, cuda, (u, d, _) = big
    ins = [8, 15]
    got = pc.threePointProject([(u, d, DIMS)], ins, MOM7)[0].cpu().numpy()
    corr = torch.empty(V, dtype=torch.complex128, device=cuda)
    for w, g in enumerate(ins):
        pc.meson_two_point(corr, u, d, V, g, 15)
        assert per_timeslice_error(got[w], torch_project(torch, corr, MOM7).cpu().numpy()) < 1e-12, g


def test_diquark_and_sequential_at_buffer_edges(big):
    torch, pc, cuda, (u, d, u2) = big
    idx = torch.from_numpy(EDGE_SITES).to(cuda)
    hu, hd, hu2 = (p.index_select(0, idx).cpu().numpy() for p in (u, d, u2))
    out = torch.empty_like(u)
    pc.baryon_diquark(out, u, d, V, 5, 5)
    got = out.index_select(0, idx).cpu().numpy()
    want = oracle.baryon_diquark(hu, hd, 5, 5)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    # sequential source into the third slot: out is overwritten, never read
    pc.baryon_sequential_two_point(u, d, out, 0, 2, V, 5, 5, 8)
    got = out.index_select(0, idx).cpu().numpy()
    want = oracle.baryon_sequential_two_point(hu, hd, hu2, 0, 2, 5, 5, 8)[2]
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    del out
    torch.cuda.empty_cache()
