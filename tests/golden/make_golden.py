"""Write golden.npz: brute-force dense-einsum vectors for every entry point.

This is an independent formulation of the reference's contractions: dense
4x4 gamma matrices built here from the DeGrand-Rossi basis, full 3x3x3
epsilon tensors, and numpy.einsum over all spin and colour indices (the style
of the reference's own cross-check, baryon_2pt.py:20-55, but with the kernel's
Gamma_mn convention Gmn[m,n] P_n[n,m] of doc/contract.md:16 and
include/baryon.cuh:55-67).  It shares no code with oracle/contract_oracle.c.

Inputs are seeded Gaussian propagators on a 2^3x4 lattice (V = 32) with the
exp(-alpha |t - t0|) profile.  Run from the repo root:

    python tests/golden/make_golden.py
"""

import os

import numpy as np

I = 1j
HERE = os.path.dirname(os.path.abspath(__file__))

G1 = np.array([[0, 0, 0, I], [0, 0, I, 0], [0, -I, 0, 0], [-I, 0, 0, 0]])
G2 = np.array([[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]], dtype=complex)
G3 = np.array([[0, 0, I, 0], [0, 0, 0, -I], [-I, 0, 0, 0], [0, I, 0, 0]])
G4 = np.array([[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]], dtype=complex)


def gamma(n):
    m = np.eye(4, dtype=complex)
    for bit, g in enumerate((G1, G2, G3, G4)):
        if (n >> bit) & 1:
            m = m @ g
    return m


G5 = gamma(15)

EPS = np.zeros((3, 3, 3))
for a in range(3):
    b, c = (a + 1) % 3, (a + 2) % 3
    EPS[a, b, c] = 1
    EPS[a, c, b] = -1

# Source-spin letter carried by (P_i, P_j, P_n) for each contraction type,
# include/contract_enum.h:1-7.
TYPES = ["IK_JL_NM", "IK_JM_NL", "IL_JK_NM", "IL_JM_NK", "IM_JK_NL", "IM_JL_NK"]
LETTERS = {t: (t[1].lower(), t[4].lower(), t[7].lower()) for t in TYPES}


def baryon_dense(pi, pj, pn, t, gij, gkl, W):
    X, Y, Z = LETTERS[TYPES[t]]
    sub = f"abc,def,ij,kl,mn,xi{X}ad,xj{Y}be,xn{Z}cf->x"
    return np.einsum(sub, EPS, EPS, gamma(gij), gamma(gkl), W, pi, pj, pn, optimize=True)


def sequential_dense(pi, pj, pn, t, s, gij, gkl, gmn):
    X, Y, Z = LETTERS[TYPES[t]]
    ops = [(f"xi{X}ad", pi), (f"xj{Y}be", pj), (f"xn{Z}cf", pn)]
    out = ops[s][0]
    rest = [o for k, o in enumerate(ops) if k != s]
    sub = "abc,def,ij,kl,mn," + ",".join(r[0] for r in rest) + "->" + out
    return np.einsum(sub, EPS, EPS, gamma(gij), gamma(gkl), gamma(gmn), *[r[1] for r in rest], optimize=True)


def as_matrix(p):
    """[V,4,4,3,3] -> [V,12,12] with row (s1,c1), column (s2,c2)."""
    return p.transpose(0, 1, 3, 2, 4).reshape(p.shape[0], 12, 12)


def meson_dense(pi, pj, gij, gkl):
    """Tr[(g5 Gij) P_i (g5 Gkl) P_j^dagger] as 12x12 matrix products."""
    A = np.kron(G5 @ gamma(gij), np.eye(3))
    B = np.kron(G5 @ gamma(gkl), np.eye(3))
    Mi, Mj = as_matrix(pi), as_matrix(pj)
    return np.einsum("ab,xbc,cd,xad->x", A, Mi, B, Mj.conj(), optimize=True)


def diquark_dense(pi, pj, gij, gkl):
    inner = np.einsum("abc,def,ij,kl,xikbe,xjmcf->xmlad", EPS, EPS, gamma(gij), gamma(gkl), pi, pj, optimize=True)
    s5 = np.diag(G5).real
    return (s5[None, :, None, None, None] * s5[None, None, :, None, None]) * inner.conj()


def propagator(dims, seed, alpha, t0=0):
    Lx, Ly, Lz, Lt = dims
    rng = np.random.Generator(np.random.PCG64(seed))
    V = Lx * Ly * Lz * Lt
    p = rng.standard_normal((V, 144)) + 1j * rng.standard_normal((V, 144))
    p = p.reshape(Lt, -1, 144) * np.exp(-alpha * np.abs(np.arange(Lt) - t0))[:, None, None]
    return p.reshape(V, 4, 4, 3, 3)


def main():
    dims = (2, 2, 2, 4)
    pi = propagator(dims, 11, 0.25)
    pj = propagator(dims, 12, 0.25)
    pn = propagator(dims, 13, 0.25)
    rng = np.random.Generator(np.random.PCG64(99))
    proj = (rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))
    out = dict(dims=np.array(dims), pi=pi, pj=pj, pn=pn, project_mn=proj)

    gpairs = [(5, 5), (0, 0), (15, 15), (1, 8), (13, 5), (10, 3)]
    for gij, gkl in gpairs:
        out[f"meson_{gij}_{gkl}"] = meson_dense(pi, pj, gij, gkl)
    for gij, gkl in [(5, 5), (13, 5), (2, 9)]:
        out[f"diquark_{gij}_{gkl}"] = diquark_dense(pi, pj, gij, gkl)
    for t in range(6):
        for gij, gkl, gmn in [(5, 5, 8), (13, 5, 0), (10, 10, 7), (3, 14, 15)]:
            out[f"baryon_{t}_{gij}_{gkl}_{gmn}"] = baryon_dense(pi, pj, pn, t, gij, gkl, gamma(gmn))
        out[f"general_{t}_5_5"] = baryon_dense(pi, pj, pn, t, 5, 5, proj)
        for s in range(3):
            out[f"sequential_{t}_{s}_5_5_8"] = sequential_dense(pi, pj, pn, t, s, 5, 5, 8)
            out[f"sequential_{t}_{s}_13_10_3"] = sequential_dense(pi, pj, pn, t, s, 13, 10, 3)
    path = os.path.join(HERE, "golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, len(out), "arrays")


if __name__ == "__main__":
    main()
