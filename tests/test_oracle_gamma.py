"""Pin the oracle's gamma matrices.

The oracle builds G(n) as a product of DeGrand-Rossi basis matrices.  Checked
against (1) the reference's own sparse tables (include/gamma.cuh:58-127,
225-271, transcribed as data in tests/golden/gamma_table.json) and (2) the
Clifford algebra, which a sign or index slip in the basis would break.
"""

import json
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
G5 = np.diag([1, 1, -1, -1]).astype(complex)


@pytest.fixture(scope="module")
def table():
    with open(os.path.join(HERE, "golden", "gamma_table.json")) as f:
        return json.load(f)


def dense_from_table(entries):
    m = np.zeros((4, 4), complex)
    for r, (c, re, im) in enumerate(entries):
        m[r, int(c)] = re + 1j * im
    return m


@pytest.mark.parametrize("g", range(16))
def test_matches_reference_tables(table, g):
    G = oracle.gamma(g)
    assert np.array_equal(dense_from_table(table["G"][g]), G)          # gamma_data<false>
    assert np.array_equal(dense_from_table(table["GT"][g]), G.T)       # gamma_data<true>
    assert np.array_equal(dense_from_table(table["Gg5"][g]), G @ G5)   # gamma_gamma5_data<false>
    assert np.array_equal(dense_from_table(table["g5GT"][g]), (G5 @ G).T)  # gamma_gamma5_data<true>


def test_clifford_algebra():
    basis = [oracle.gamma(1 << mu) for mu in range(4)]
    for mu in range(4):
        for nu in range(4):
            anti = basis[mu] @ basis[nu] + basis[nu] @ basis[mu]
            assert np.allclose(anti, 2 * np.eye(4) * (mu == nu))
        assert np.allclose(basis[mu], basis[mu].conj().T)  # hermitian
    assert np.array_equal(oracle.gamma(15), G5)
    assert np.allclose(basis[0] @ basis[1] @ basis[2] @ basis[3], G5)


def test_one_nonzero_per_row_symmetric_pattern():
    for g in range(16):
        G = oracle.gamma(g)
        nz = G != 0
        assert (nz.sum(axis=1) == 1).all()
        assert np.array_equal(nz, nz.T)
        # C gamma_5 = G(5), C = gamma_2 gamma_4 = G(10) (baryon_2pt.py:79-81)
    assert np.allclose(oracle.gamma(2) @ oracle.gamma(8), oracle.gamma(10))
    assert np.allclose(oracle.gamma(10) @ G5, oracle.gamma(5))
