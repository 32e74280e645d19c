"""CPU FP64 oracle for the propagator-contraction path (test infrastructure).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package; the product library never does.  The
numerics live in ``oracle/contract_oracle.c`` (plain C99 + OpenMP, built by
``make -C oracle``); this module is a thin numpy/ctypes wrapper around it.

Array conventions follow the reference ABI (include/contract.h:15-27):
propagators are complex128 ``[V, 4, 4, 3, 3]`` =
``[site][sink spin][source spin][sink colour][source colour]``
(include/load_store.cuh:99-109); site correlators are complex128 ``[V]``.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libcontract_ref.so")

IK_JL_NM, IK_JM_NL, IL_JK_NM, IL_JM_NK, IM_JK_NL, IM_JL_NK = range(6)
SEQUENTIAL_I, SEQUENTIAL_J, SEQUENTIAL_N = range(3)

_lib = None


def build(force: bool = False) -> str:
    """Compile the C oracle with gcc (make -C oracle)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        S = ctypes.c_size_t
        I = ctypes.c_int
        L.oracle_gamma.argtypes = [I, P]
        L.oracle_meson_two_point.argtypes = [P, P, P, S, I, I]
        L.oracle_meson_all_source_two_point.argtypes = [P, P, P, S, I]
        L.oracle_meson_all_sink_two_point.argtypes = [P, P, P, S, I]
        L.oracle_baryon_diquark.argtypes = [P, P, P, S, I, I]
        L.oracle_baryon_two_point.argtypes = [P, P, P, P, I, S, I, I, I]
        L.oracle_baryon_general_two_point.argtypes = [P, P, P, P, I, S, I, I, P]
        L.oracle_baryon_sequential_two_point.argtypes = [P, P, P, I, I, S, I, I, I]
        L.oracle_site_coords.argtypes = [S, P, I, P]
        L.oracle_project.argtypes = [P, P, I, P, I, P, I, P, P]
        L.oracle_baryon_twopoint_project.argtypes = [P, P, P, P, P, P, I, I, I, P, I, P, I, P, I, P, P]
        L.oracle_threepoint_project.argtypes = [P, P, P, P, I, P, I, P, I, P, P]
        L.oracle_num_threads.restype = I
        L.oracle_set_num_threads.argtypes = [I]
        L.oracle_set_num_threads.restype = None
        _lib = L
    return _lib


def _c128(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _prop(p):
    p = _c128(p)
    if p.ndim == 5:
        assert p.shape[1:] == (4, 4, 3, 3), p.shape
    else:
        assert p.size % 144 == 0
    return p


def _volume(p):
    return p.size // 144


def num_threads() -> int:
    return lib().oracle_num_threads()


def set_num_threads(n: int) -> None:
    """Threads of the following oracle calls (n <= 0: the OpenMP default)."""
    lib().oracle_set_num_threads(int(n))


def gamma(n: int) -> np.ndarray:
    """Dense DeGrand-Rossi G(n), 4x4 complex128 (include/gamma.cuh:176-223)."""
    out = np.zeros((4, 4), np.complex128)
    lib().oracle_gamma(n, _ptr(out))
    return out


def meson_two_point(propag_i, propag_j, gamma_ij: int, gamma_kl: int) -> np.ndarray:
    """include/contract.h:16 -> include/meson.cuh:32-50."""
    pi, pj = _prop(propag_i), _prop(propag_j)
    V = _volume(pi)
    out = np.zeros(V, np.complex128)
    lib().oracle_meson_two_point(_ptr(out), _ptr(pi), _ptr(pj), V, gamma_ij, gamma_kl)
    return out


def meson_all_source_two_point(propag_i, propag_j, gamma_ij: int) -> np.ndarray:
    """include/contract.h:17 -> include/meson_all.cuh:52-73; returns [16, V]."""
    pi, pj = _prop(propag_i), _prop(propag_j)
    V = _volume(pi)
    out = np.zeros((16, V), np.complex128)
    lib().oracle_meson_all_source_two_point(_ptr(out), _ptr(pi), _ptr(pj), V, gamma_ij)
    return out


def meson_all_sink_two_point(propag_i, propag_j, gamma_kl: int) -> np.ndarray:
    """include/contract.h:18 -> include/meson_all.cuh:82-103; returns [16, V]."""
    pi, pj = _prop(propag_i), _prop(propag_j)
    V = _volume(pi)
    out = np.zeros((16, V), np.complex128)
    lib().oracle_meson_all_sink_two_point(_ptr(out), _ptr(pi), _ptr(pj), V, gamma_kl)
    return out


def baryon_diquark(propag_i, propag_j, gamma_ij: int, gamma_kl: int) -> np.ndarray:
    """include/contract.h:19 -> include/diquark.cuh:32-58; returns [V,4,4,3,3]."""
    pi, pj = _prop(propag_i), _prop(propag_j)
    V = _volume(pi)
    out = np.zeros((V, 4, 4, 3, 3), np.complex128)
    lib().oracle_baryon_diquark(_ptr(out), _ptr(pi), _ptr(pj), V, gamma_ij, gamma_kl)
    return out


def baryon_two_point(propag_i, propag_j, propag_n, contract_type: int, gamma_ij: int, gamma_kl: int,
                     gamma_mn: int) -> np.ndarray:
    """include/contract.h:20-21 -> include/baryon.cuh:34-105."""
    pi, pj, pn = _prop(propag_i), _prop(propag_j), _prop(propag_n)
    V = _volume(pi)
    out = np.zeros(V, np.complex128)
    lib().oracle_baryon_two_point(_ptr(out), _ptr(pi), _ptr(pj), _ptr(pn), int(contract_type), V, gamma_ij,
                                  gamma_kl, gamma_mn)
    return out


def baryon_general_two_point(propag_i, propag_j, propag_n, contract_type: int, gamma_ij: int, gamma_kl: int,
                             project_mn) -> np.ndarray:
    """include/contract.h:22-24 -> include/baryon_general.cuh:36-112."""
    pi, pj, pn = _prop(propag_i), _prop(propag_j), _prop(propag_n)
    proj = _c128(project_mn).reshape(16)
    V = _volume(pi)
    out = np.zeros(V, np.complex128)
    lib().oracle_baryon_general_two_point(_ptr(out), _ptr(pi), _ptr(pj), _ptr(pn), int(contract_type), V, gamma_ij,
                                          gamma_kl, _ptr(proj))
    return out


def baryon_sequential_two_point(propag_i, propag_j, propag_n, contract_type: int, sequential_type: int,
                                gamma_ij: int, gamma_kl: int, gamma_mn: int):
    """include/contract.h:25-27 -> include/baryon_sequential.cuh:36-267.

    Works on copies and returns (propag_i, propag_j, propag_n) after the call;
    the propagator named by ``sequential_type`` holds dC/dP_X.
    """
    pi, pj, pn = _prop(propag_i).copy(), _prop(propag_j).copy(), _prop(propag_n).copy()
    V = _volume(pi)
    lib().oracle_baryon_sequential_two_point(_ptr(pi), _ptr(pj), _ptr(pn), int(contract_type), int(sequential_type),
                                             V, gamma_ij, gamma_kl, gamma_mn)
    return pi, pj, pn


def site_coords(V_dims, order: int = 0) -> np.ndarray:
    """[V, 4] integer (x, y, z, t) of every site for the given site order."""
    dims = np.ascontiguousarray(V_dims, dtype=np.int32)
    V = int(np.prod(dims))
    out = np.zeros((V, 4), np.int32)
    c = np.zeros(4, np.int32)
    L = lib()
    for s in range(V):
        L.oracle_site_coords(s, _ptr(dims), order, _ptr(c))
        out[s] = c
    return out


def _geom(v):
    return None if v is None else np.ascontiguousarray(list(v)[:3], dtype=np.int32)


def project(site_corr, dims, momenta, order: int = 0, global_dims=None, origin=None) -> np.ndarray:
    """Timeslice + momentum projection: site_corr [V, nq] -> [nq, np, Lt].
    global_dims / origin: the local lattice as part of a larger one (phases in
    global coordinates, momenta in units of 2pi / global extent)."""
    c = _c128(site_corr)
    if c.ndim == 1:
        c = c[:, None].copy()
    nq = c.shape[1]
    d = np.ascontiguousarray(dims, dtype=np.int32)
    mom = np.ascontiguousarray(momenta, dtype=np.int32).reshape(-1, 3)
    out = np.zeros((nq, mom.shape[0], int(d[3])), np.complex128)
    gl, og = _geom(global_dims), _geom(origin)
    lib().oracle_project(_ptr(out), _ptr(c), nq, _ptr(d), order, _ptr(mom), mom.shape[0],
                         None if gl is None else _ptr(gl), None if og is None else _ptr(og))
    return out


def baryon_twopoint_project(propag_i, propag_j, propag_n, terms, gamma_ij: int, gamma_kl: int, insertions, dims,
                            momenta, order: int = 0, global_dims=None, origin=None) -> np.ndarray:
    """Fused workload: sum_terms w * baryon_general_two_point(type, W) for each
    insertion W (a gamma index g -> G(g), or a 4x4 projector), projected to
    [len(insertions), len(momenta), Lt]."""
    pi, pj, pn = _prop(propag_i), _prop(propag_j), _prop(propag_n)
    types = np.ascontiguousarray([t for t, _ in terms], dtype=np.int32)
    w = np.ascontiguousarray([complex(x) for _, x in terms], dtype=np.complex128)
    W = np.ascontiguousarray(np.stack([gamma(int(g)) if np.ndim(g) == 0 else np.asarray(g, np.complex128)
                                       for g in insertions]))
    d = np.ascontiguousarray(dims, dtype=np.int32)
    mom = np.ascontiguousarray(momenta, dtype=np.int32).reshape(-1, 3)
    out = np.zeros((W.shape[0], mom.shape[0], int(d[3])), np.complex128)
    gl, og = _geom(global_dims), _geom(origin)
    lib().oracle_baryon_twopoint_project(_ptr(out), _ptr(pi), _ptr(pj), _ptr(pn), _ptr(types), _ptr(w), len(types),
                                         gamma_ij, gamma_kl, _ptr(W), W.shape[0], _ptr(d), order, _ptr(mom),
                                         mom.shape[0], None if gl is None else _ptr(gl),
                                         None if og is None else _ptr(og))
    return out


def threepoint_project(fwd, seq, insertions, dims, momenta, order: int = 0, global_dims=None,
                       origin=None) -> np.ndarray:
    """Three-point step: out[w][p][t] = projection of
    Tr_s Tr_c[g5 Gamma_w F(x) Sigma(x)^dagger] for the forward propagator F,
    the sequential propagator Sigma and each insertion Gamma_w (a gamma index
    g -> G(g), or a dense 4x4 matrix)."""
    f, s = _prop(fwd), _prop(seq)
    G = np.ascontiguousarray(np.stack([gamma(int(g)) if np.ndim(g) == 0 else np.asarray(g, np.complex128)
                                       for g in insertions]))
    d = np.ascontiguousarray(dims, dtype=np.int32)
    mom = np.ascontiguousarray(momenta, dtype=np.int32).reshape(-1, 3)
    out = np.zeros((G.shape[0], mom.shape[0], int(d[3])), np.complex128)
    gl, og = _geom(global_dims), _geom(origin)
    lib().oracle_threepoint_project(_ptr(out), _ptr(f), _ptr(s), _ptr(G), G.shape[0], _ptr(d), order, _ptr(mom),
                                    mom.shape[0], None if gl is None else _ptr(gl), None if og is None else _ptr(og))
    return out


def sequential_propagator_unit_dirac(seq_source) -> np.ndarray:
    """The sequential propagator Sigma = D^-1 sigma for the unit Dirac
    operator D = 1, where sigma(y)[s1][s2](c1,c2) = g5[s1] conj(seq(y)[s1][s2](c1,c2))
    is the source a solver is given for the sequential source seq = dC/dP_X
    (g5-hermiticity S(y,x) = g5 S(x,y)^dagger g5 turns the backward line into
    a forward solve).  With D = 1 the three-point function of insertion Gamma
    equals the two-point function with P_X replaced by Gamma P_X."""
    s = _prop(seq_source)
    g5 = np.array([1.0, 1.0, -1.0, -1.0])
    return np.ascontiguousarray(g5[None, :, None, None, None] * np.conj(s))


def meson_grid_project(propag_i, propag_j, dims, momenta, order: int = 0, global_dims=None, origin=None) -> np.ndarray:
    """The batched meson grid of one request: out[g1][g2][p][t] = projection
    of meson_two_point(propag_i, propag_j; gamma_ij = g1, gamma_kl = g2)
    (include/contract.h:16, include/meson.cuh:32-50), one oracle call per
    gamma pair, then the oracle's timeslice/momentum projection."""
    pi, pj = _prop(propag_i), _prop(propag_j)
    V = _volume(pi)
    corr = np.empty((V, 256), np.complex128)
    for g1 in range(16):
        for g2 in range(16):
            corr[:, g1 * 16 + g2] = meson_two_point(pi, pj, g1, g2)
    out = project(corr, dims, momenta, order, global_dims, origin)
    return out.reshape(16, 16, out.shape[1], out.shape[2])


# ---------------------------------------------------------------------------
# Synthetic propagators
# ---------------------------------------------------------------------------

def synthetic_propagator(dims, seed: int, alpha: float, t0: int = 0, dtype=np.complex128) -> np.ndarray:
    """Propagator with iid N(0,1) real and imaginary parts scaled per timeslice
    by exp(-alpha |t - t0|), lexicographic [Lt][Lz][Ly][Lx][4][4][3][3]."""
    Lx, Ly, Lz, Lt = (int(v) for v in dims)
    rng = np.random.Generator(np.random.PCG64(seed))
    V = Lx * Ly * Lz * Lt
    re = rng.standard_normal((V, 144))
    im = rng.standard_normal((V, 144))
    p = (re + 1j * im).reshape(Lt, Lx * Ly * Lz, 144)
    t = np.arange(Lt)
    p *= np.exp(-alpha * np.abs(t - t0))[:, None, None]
    return p.reshape(V, 4, 4, 3, 3).astype(dtype)


def ref_library_path() -> str:
    """Path of the reference's own HIP build (oracle/_ref), if it was built."""
    return _REF_PATH
