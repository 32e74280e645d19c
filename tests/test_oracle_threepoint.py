"""Pins of the oracle's three-point step (oracle_threepoint_project), no GPU.

* An independent numpy einsum of Tr_s Tr_c[g5 Gamma F Sigma^dagger] with the
  oracle's own projection.
* Per gamma index it equals meson_two_point(F, Sigma; Gamma, 15), the
  reference's meson contraction (include/meson.cuh:32-50, pinned against the
  reference library in test_oracle_ref_vectors.py).
* The unit-Dirac chain: with D = 1 the sequential propagator is the solver
  source sigma = g5 conj(seq_X) of the reference's sequential source
  (include/baryon_sequential.cuh:165-267), and the three-point function of
  Gamma equals the two-point function with P_X -> Gamma P_X, for all six
  contraction types and all three targets.  A dropped g5, a missing
  conjugate or a transposed insertion fails it."""

import numpy as np
import pytest

import oracle

DIMS = (2, 2, 2, 4)
MOMS = [(0, 0, 0), (1, 0, 0), (-1, 1, 0)]
G5 = np.diag([1.0, 1.0, -1.0, -1.0])


def props(n=3, alpha=0.3):
    return [oracle.synthetic_propagator(DIMS, 500 + k, alpha, t0=k) for k in range(n)]


def test_matches_independent_einsum():
    f, s = props(2)
    dense = np.array([[0.5, 1j, 0, 0], [0, -1, 0, 2], [0.25, 0, 1, 0], [0, 0, -1j, 3]])
    ins = [0, 5, 9, 15, dense]
    mats = [oracle.gamma(g) if np.ndim(g) == 0 else dense for g in ins]
    site = np.stack([np.einsum("ab,xbscd,xascd->x", G5 @ m, f, np.conj(s)) for m in mats], axis=1)
    want = oracle.project(site, DIMS, MOMS)
    got = oracle.threepoint_project(f, s, ins, DIMS, MOMS)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


def test_equals_meson_with_g5_at_the_source():
    f, s = props(2)
    got = oracle.threepoint_project(f, s, list(range(16)), DIMS, MOMS)
    site = np.stack([oracle.meson_two_point(f, s, g, 15) for g in range(16)], axis=1)
    want = oracle.project(site, DIMS, MOMS)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


@pytest.mark.parametrize("t", range(6))
@pytest.mark.parametrize("x", range(3))
def test_unit_dirac_chain(t, x):
    ps = props(3)
    seq = oracle.baryon_sequential_two_point(*ps, t, x, 5, 5, 8)[x]
    sigma = oracle.sequential_propagator_unit_dirac(seq)
    ins = [0, 1, 8, 14, 15]
    got = oracle.threepoint_project(ps[x], sigma, ins, DIMS, MOMS)
    for w, g in enumerate(ins):
        slots = [p.copy() for p in ps]
        slots[x] = np.einsum("ab,xbscd->xascd", oracle.gamma(g), ps[x])
        want = oracle.baryon_twopoint_project(*slots, [(t, 1.0)], 5, 5, [8], DIMS, MOMS)
        assert np.abs(got[w] - want[0]).max() <= 1e-12 * np.abs(want).max(), (t, x, g)


def test_chain_fails_without_the_g5_factor():
    """The chain is sensitive to the source convention: leaving out g5 (or the
    conjugate) breaks it for an insertion that mixes upper and lower spins."""
    ps = props(3)
    seq = oracle.baryon_sequential_two_point(*ps, 0, 2, 5, 5, 8)[2]
    wrong = [np.conj(seq), seq * G5.diagonal()[None, :, None, None, None]]
    slots = [p.copy() for p in ps]
    slots[2] = np.einsum("ab,xbscd->xascd", oracle.gamma(8), ps[2])
    want = oracle.baryon_twopoint_project(*slots, [(0, 1.0)], 5, 5, [8], DIMS, MOMS)
    for sigma in wrong:
        got = oracle.threepoint_project(ps[2], sigma, [8], DIMS, MOMS)
        assert np.abs(got - want).max() > 1e-3 * np.abs(want).max()
