"""Identities that tie the oracle's entry points to each other and to closed
forms, so that a dropped term, a wrong sign or a transposed operand in any
one of them is caught by a different formulation."""

import numpy as np
import pytest

import oracle

G5 = np.diag([1, 1, -1, -1]).astype(complex)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def props():
    dims = (2, 2, 2, 4)
    return [oracle.synthetic_propagator(dims, s, 0.25) for s in (1, 2, 3)]


def test_general_with_gamma_equals_sparse(props):
    """baryon_general_two_point(project = G(g)) == baryon_two_point(g)
    (include/baryon_general.cuh:62-67 vs include/baryon.cuh:55-67)."""
    pi, pj, pn = props
    for t in range(6):
        for g in (0, 4, 8, 13):
            a = oracle.baryon_two_point(pi, pj, pn, t, 5, 10, g)
            b = oracle.baryon_general_two_point(pi, pj, pn, t, 5, 10, oracle.gamma(g))
            assert rel(b, a) < 1e-14


def test_general_is_linear_in_projector(props):
    pi, pj, pn = props
    rng = np.random.default_rng(5)
    P1 = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    P2 = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    for t in range(6):
        a = oracle.baryon_general_two_point(pi, pj, pn, t, 5, 5, P1 + 2j * P2)
        b = oracle.baryon_general_two_point(pi, pj, pn, t, 5, 5, P1)
        c = oracle.baryon_general_two_point(pi, pj, pn, t, 5, 5, P2)
        assert rel(a, b + 2j * c) < 1e-13


def test_diquark_meson_is_baryon(props):
    """baryon(IK_JM_NL; Pi,Pj,Pn; Gij,Gkl,Gmn) == meson(Pn, diquark(Pi,Pj;Gij,Gkl); Gmn, 0)."""
    pi, pj, pn = props
    for gij, gkl, gmn in [(5, 5, 8), (13, 10, 3), (2, 7, 15), (0, 9, 6)]:
        b = oracle.baryon_two_point(pi, pj, pn, 1, gij, gkl, gmn)
        d = oracle.baryon_diquark(pi, pj, gij, gkl)
        m = oracle.meson_two_point(pn, d, gmn, 0)
        assert rel(m, b) < 1e-13


@pytest.mark.parametrize("t", range(6))
@pytest.mark.parametrize("s", range(3))
def test_sequential_contracts_back_to_two_point(props, t, s):
    """C is linear in each propagator, so sum_{site entries} P_X * dC/dP_X = C."""
    pi, pj, pn = props
    for gij, gkl, gmn in [(5, 5, 8), (13, 10, 3)]:
        C = oracle.baryon_two_point(pi, pj, pn, t, gij, gkl, gmn)
        seq = oracle.baryon_sequential_two_point(pi, pj, pn, t, s, gij, gkl, gmn)[s]
        PX = (pi, pj, pn)[s]
        back = (PX.reshape(len(C), -1) * seq.reshape(len(C), -1)).sum(axis=1)
        assert rel(back, C) < 1e-13


def test_sequential_never_reads_target(props):
    pi, pj, pn = props
    for s in range(3):
        a = oracle.baryon_sequential_two_point(pi, pj, pn, 3, s, 5, 5, 8)[s]
        junk = [p.copy() for p in props]
        junk[s][:] = 7.0 + 3.0j
        b = oracle.baryon_sequential_two_point(*junk, 3, s, 5, 5, 8)[s]
        assert np.array_equal(a, b)
        # and the other two are untouched
        outs = oracle.baryon_sequential_two_point(pi, pj, pn, 3, s, 5, 5, 8)
        for o in range(3):
            if o != s:
                assert np.array_equal(outs[o], props[o])


def test_sequential_invalid_type_is_noop(props):
    outs = oracle.baryon_sequential_two_point(*props, 0, 3, 5, 5, 8)
    for a, b in zip(outs, props):
        assert np.array_equal(a, b)


def test_meson_all_is_sixteen_mesons(props):
    pi, pj, _ = props
    src = oracle.meson_all_source_two_point(pi, pj, 5)
    snk = oracle.meson_all_sink_two_point(pi, pj, 7)
    for g in range(16):
        assert rel(src[g], oracle.meson_two_point(pi, pj, 5, g)) == 0
        assert rel(snk[g], oracle.meson_two_point(pi, pj, g, 7)) == 0


def test_meson_cfg0_trace_relation(props):
    """tr[G S1 G' g5 S2^+ g5] = (-1)^popcount(G') meson(S1, S2, G, G')."""
    s1, s2, _ = props

    def mat(p):
        return p.transpose(0, 1, 3, 2, 4).reshape(p.shape[0], 12, 12)

    M1, M2 = mat(s1), mat(s2)
    E = np.eye(3)
    for g in range(16):
        A = np.kron(oracle.gamma(g), E)
        B = np.kron(oracle.gamma(g) @ G5, E)
        g5 = np.kron(G5, E)
        direct = np.einsum("ab,xbc,cd,xed,ea->x", A, M1, B, M2.conj(), g5, optimize=True)
        sign = (-1) ** bin(g).count("1")
        assert rel(sign * oracle.meson_two_point(s1, s2, g, g), direct) < 1e-13


def test_baryon_sum_over_colour_is_symmetric(props):
    """ee(X,Y,Z) is totally symmetric: permuting (P_i, P_j, P_n) together with
    their spin roles leaves C unchanged (IK_JL_NM with Gij symmetric)."""
    pi, pj, pn = props
    # Gij = Gkl = C g5 = G(5) is antisymmetric; swapping i<->j and k<->l flips
    # both factors: C(IK_JL_NM; Pi, Pj) = C(IK_JL_NM; Pj, Pi).
    a = oracle.baryon_two_point(pi, pj, pn, 0, 5, 5, 8)
    b = oracle.baryon_two_point(pj, pi, pn, 0, 5, 5, 8)
    assert rel(b, a) < 1e-13


# ---------------------------------------------------------------------------
# projection
# ---------------------------------------------------------------------------

MOMENTA = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, -1), (2, -1, 1)]


def test_projection_matches_fft():
    dims = (4, 6, 2, 3)
    Lx, Ly, Lz, Lt = dims
    V = Lx * Ly * Lz * Lt
    rng = np.random.default_rng(11)
    c = rng.standard_normal((V, 2)) + 1j * rng.standard_normal((V, 2))
    out = oracle.project(c, dims, MOMENTA)
    for q in range(2):
        F = np.fft.fftn(c[:, q].reshape(Lt, Lz, Ly, Lx), axes=(1, 2, 3))
        for ip, (px, py, pz) in enumerate(MOMENTA):
            ref = F[:, pz % Lz, py % Ly, px % Lx]
            assert np.abs(out[q, ip] - ref).max() < 1e-12 * np.abs(F).max()


def test_projection_plane_wave_and_uniform():
    dims = (4, 4, 4, 2)
    co = oracle.site_coords(dims, 0)
    q = np.array([1, -1, 0])
    ph = np.exp(2j * np.pi * (co[:, :3] @ q) / 4.0)
    out = oracle.project(np.stack([np.ones(len(co)), ph], axis=1), dims, MOMENTA)
    V3 = 64
    for ip, p in enumerate(MOMENTA):
        exp_u = V3 if p == (0, 0, 0) else 0
        exp_w = V3 if tuple(np.mod(p, 4)) == tuple(np.mod(q, 4)) else 0
        assert np.allclose(out[0, ip], exp_u, atol=1e-12)
        assert np.allclose(out[1, ip], exp_w, atol=1e-11)


def test_even_odd_coordinates():
    dims = (4, 2, 2, 4)
    Lx, Ly, Lz, Lt = dims
    co = oracle.site_coords(dims, 1)
    V = len(co)
    seen = set()
    for s, (x, y, z, t) in enumerate(co):
        par = (x + y + z + t) % 2
        expect = par * V // 2 + ((t * Lz + z) * Ly + y) * Lx // 2 + x // 2
        assert expect == s
        seen.add((x, y, z, t))
    assert len(seen) == V


def test_fused_twopoint_is_projected_sum(props):
    dims = (2, 2, 2, 4)
    pi, pj, pn = props
    terms = [(0, 1.0), (5, -0.5 + 0.25j)]
    gammas = [0, 8, 15]
    mom = [(0, 0, 0), (1, 0, 0), (0, 1, 1)]
    out = oracle.baryon_twopoint_project(pi, pj, pn, terms, 5, 5, gammas, dims, mom)
    Lx, Ly, Lz, Lt = dims
    for ig, g in enumerate(gammas):
        site = sum(w * oracle.baryon_two_point(pi, pj, pn, t, 5, 5, g) for t, w in terms)
        F = np.fft.fftn(site.reshape(Lt, Lz, Ly, Lx), axes=(1, 2, 3))
        for ip, (px, py, pz) in enumerate(mom):
            assert np.abs(out[ig, ip] - F[:, pz % Lz, py % Ly, px % Lx]).max() < 1e-12 * np.abs(F).max()
