"""The reference's own consistency check, baryon_2pt.py:20-55, restated in
numpy: its opt_einsum form of baryonTwoPoint, written with the same subscript
strings, against the oracle's baryon_general_two_point.

The demo's einsum contracts Gamma_mn[m, n] with P_m[m, n] while the kernels
(include/baryon_general.cuh:62-84) and the oracle contract Gamma_mn[m, n] with
P[n, m]: the two agree for a symmetric Gamma_mn -- the demo's projector
P = (1 - gamma_4) / 2 -- and for any other one after transposing it.  The
demo's enum names (IK_JL_MN, ...) are the aliases of contract_enum.h:14-19
that rename m <-> n.  Gammas as in baryon_2pt.py:72-81: C = gamma_2 gamma_4,
CG_A = C gamma_4 gamma_5, CG_B = C gamma_5."""

import numpy as np
import pytest

import oracle

EPS = np.zeros((3, 3, 3))
for _i in range(3):
    EPS[_i, (_i + 1) % 3, (_i + 2) % 3] = 1
    EPS[_i, (_i + 2) % 3, (_i + 1) % 3] = -1

# baryon_2pt.py:31-42, per type of contract_enum.h (value = BaryonContractType)
DEMO = {
    0: "ikad,jlbe,mncf",  # IK_JL_MN
    1: "ikad,jnbe,mlcf",  # IK_JN_ML
    2: "ilad,jkbe,mncf",  # IL_JK_MN
    3: "ilad,jnbe,mkcf",  # IL_JN_MK
    4: "inad,jkbe,mlcf",  # IN_JK_ML
    5: "inad,jlbe,mkcf",  # IN_JL_MK
}


def demo_einsum(pi, pj, pm, t, gij, gkl, gmn):
    a, b, c = DEMO[t].split(",")
    sub = f"abc,def,ij,kl,mn,x{a},x{b},x{c}->x"
    P = [np.asarray(p).reshape(-1, 4, 4, 3, 3) for p in (pi, pj, pm)]
    return np.einsum(sub, EPS, EPS, gij, gkl, gmn, *P, optimize=True)


@pytest.fixture(scope="module")
def props():
    dims = (2, 2, 2, 2)
    return [oracle.synthetic_propagator(dims, s, 0.3) for s in (11, 12, 13)]


def demo_gammas():
    g = oracle.gamma
    C = g(2) @ g(8)
    return C @ g(8) @ g(15), C @ g(15), (g(0) - g(8)) / 2


def test_demo_gammas_are_the_documented_indices():
    cga, cgb, P = demo_gammas()
    assert np.allclose(cga, -oracle.gamma(13)) and np.allclose(cgb, oracle.gamma(5))
    assert np.allclose(P, P.T)


@pytest.mark.parametrize("t", range(6))
def test_demo_check_holds_for_the_demo_projector(props, t):
    """baryon_2pt.py compares einsum and pycontract for IN_JK_ML and IK_JL_MN
    with CG_A, CG_B, P; restated here for all six types."""
    cga, cgb, P = demo_gammas()
    want = demo_einsum(*props, t, cga, cgb, P)
    got = -oracle.baryon_general_two_point(*props, t, 13, 5, P)  # CG_A = -G(13)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("t", range(6))
def test_nonsymmetric_projector_enters_transposed(props, t):
    rng = np.random.default_rng(t)
    W = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    want = demo_einsum(*props, t, oracle.gamma(5), oracle.gamma(10), W)
    got = oracle.baryon_general_two_point(*props, t, 5, 10, W.T)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    other = oracle.baryon_general_two_point(*props, t, 5, 10, W)
    assert np.abs(other - want).max() > 1e-3 * np.abs(want).max()
