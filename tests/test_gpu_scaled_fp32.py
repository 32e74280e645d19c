"""FP32 site-scaled storage (CONTRACT_FP32_SCALED) on the GPU.

The converter is checked against numpy's exact reading of the format: e(x) =
floor(log2 max |Re|, |Im| of site x) (frexp), stored entries = complex64 of
P(x) 2^-e(x) (scaling by a power of two is exact, so the only rounding is
FP64 -> FP32).  The fused contraction on scaled propagators is checked against
the FP64 oracle on the original complex128 propagators at 1e-6 per-timeslice
relative error (the FP32 storage rounding is ~6e-8 per entry), including a
64-timeslice exp(-0.8|t|) decay whose three-propagator products reach
~1e-66, far below FP32's range."""

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


def test_converter_matches_the_format(dev):
    torch, pc, cuda = dev
    V = 4 * 4 * 2 * 3 + 2  # 98 sites: ragged last tile
    rng = np.random.default_rng(5)
    p = (rng.standard_normal((V, 144)) + 1j * rng.standard_normal((V, 144))).astype(np.complex128)
    p *= np.exp(rng.uniform(-80, 80, V))[:, None]  # sites from ~1e-35 to ~1e35
    p[3] = 0.0                                     # an all-zero site
    p[7] = 1e-300                                  # below the clamp: e = -127
    p[11] /= np.abs(p[11]).max()
    p[11, 5] = 2.0 ** 40                           # one dominant entry
    s = pc.toScaledFP32(torch.from_numpy(p.reshape(V, 4, 4, 3, 3)).to(cuda))
    assert s.data.numel() == V * 1153 == pc.scaled_fp32_bytes(V)
    e = s.exponents().cpu().numpy().astype(int)
    mx = np.maximum(np.abs(p.real), np.abs(p.imag)).max(axis=1)
    want_e = np.where(mx > 0, np.frexp(mx)[1] - 1, 0).clip(-127, 127)
    assert np.array_equal(e, want_e)
    assert e[3] == 0 and e[7] == -127 and e[11] == 40
    want = (p * np.ldexp(1.0, -want_e)[:, None]).astype(np.complex64)
    got = s.entries().cpu().numpy().reshape(V, 144)
    assert np.array_equal(got, want)
    smx = np.maximum(np.abs(got.real), np.abs(got.imag)).max(axis=1)
    ok = (mx > 0) & (want_e > -127)
    assert np.all((smx[ok] >= 1.0) & (smx[ok] < 2.0))


def run_scaled(dev, dims, order, alpha, gammas, moms, seeds=(41, 42, 43), terms=PROTON, gij=5, gkl=5):
    torch, pc, cuda = dev
    ps = [oracle.synthetic_propagator(dims, s, alpha, t0=1) for s in seeds]
    sc = [pc.toScaledFP32(torch.from_numpy(p).to(cuda)) for p in ps]
    got = pc.baryonTwoPointProject(*sc, terms, gij, gkl, gammas, moms, dims, site_order=order).cpu().numpy()
    want = oracle.baryon_twopoint_project(*ps, terms, gij, gkl, gammas, dims, moms, order=order)
    return got, want


@pytest.mark.parametrize("order", [0, 1])
def test_proton_scaled_matches_fp64_oracle(dev, order):
    got, want = run_scaled(dev, (4, 4, 4, 6), order, 0.25, list(range(16)), MOM27)
    assert per_timeslice_error(got, want) < 1e-6


def test_scaled_steep_decay_keeps_every_timeslice(dev):
    got, want = run_scaled(dev, (4, 4, 4, 64), 0, 0.8, list(range(16)), [(0, 0, 0), (1, 0, 0)])
    assert np.abs(want[:, 0, -1]).max() < 1e-55
    assert per_timeslice_error(got, want) < 1e-6


def test_scaled_runtime_gamma_and_ragged(dev):
    got, want = run_scaled(dev, (3, 3, 3, 5), 0, 0.3, [0, 7, 15], MOM27[:7], terms=[(2, 1.0), (4, 0.5j)], gij=13,
                           gkl=10)
    assert per_timeslice_error(got, want) < 1e-6


def test_scaled_batch_equals_single_calls(dev):
    torch, pc, cuda = dev
    cases = [((4, 4, 4, 8), 0), ((6, 4, 2, 6), 1), ((3, 3, 3, 5), 0)]
    reqs = []
    for n, (dims, order) in enumerate(cases):
        ps = [oracle.synthetic_propagator(dims, 300 + 3 * n + k, 0.25) for k in range(3)]
        reqs.append((*[pc.toScaledFP32(torch.from_numpy(p).to(cuda)) for p in ps], dims, order))
    outs = pc.baryonTwoPointProjectBatch(reqs, PROTON, 5, 5, [0, 8], MOM27)
    for (pi, pj, pn, dims, order), got in zip(reqs, outs):
        alone = pc.baryonTwoPointProject(pi, pj, pn, PROTON, 5, 5, [0, 8], MOM27, dims, site_order=order)
        assert torch.equal(got, alone)


def test_mixed_storage_forms_are_rejected(dev):
    torch, pc, cuda = dev
    dims = (2, 2, 2, 2)
    p = torch.from_numpy(oracle.synthetic_propagator(dims, 1, 0.1)).to(cuda)
    with pytest.raises(TypeError):
        pc.baryonTwoPointProject(pc.toScaledFP32(p), p, p, PROTON, 5, 5, [0], [(0, 0, 0)], dims)


@pytest.mark.parametrize("scaled", [False, True])
def test_nucleon_kernel_equals_runtime_terms(dev, scaled):
    """(IK_JL_NM, IM_JL_NK) with Cg5 takes the compile-time nucleon kernel in
    FP32; the same terms in the other order take the runtime-dispatch kernel.
    Both sum two FP32 tails in FP64, which is exact, so they agree to 1e-12."""
    torch, pc, cuda = dev
    dims = (4, 4, 2, 6)
    ps = [oracle.synthetic_propagator(dims, s, 0.3, t0=2) for s in (61, 62, 63)]
    d = [torch.from_numpy(p).to(cuda) for p in ps]
    d = [pc.toScaledFP32(p) for p in d] if scaled else [p.to(torch.complex64) for p in d]
    a = pc.baryonTwoPointProject(*d, [(0, 1.0), (5, 0.5)], 5, 5, list(range(16)), MOM27, dims).cpu().numpy()
    b = pc.baryonTwoPointProject(*d, [(5, 0.5), (0, 1.0)], 5, 5, list(range(16)), MOM27, dims).cpu().numpy()
    assert per_timeslice_error(a, b) < 1e-12
    want = oracle.baryon_twopoint_project(*ps, [(0, 1.0), (5, 0.5)], 5, 5, list(range(16)), dims, MOM27)
    assert per_timeslice_error(a, want) < 1e-6
