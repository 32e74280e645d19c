"""Projection of a sub-lattice in global coordinates (one rank's part of an
MPI grid): splitting a lattice into parts along x (or t), projecting each part
with its origin and the global extents, and adding (or stacking) the parts
gives the projection of the whole lattice.  Pins the oracle's global-geometry
projection against its whole-lattice one."""

import numpy as np
import pytest

import oracle

MOM = [(0, 0, 0), (1, 0, 0), (-1, 2, 0), (0, -1, 1), (2, 1, -1)]


def split(arr, dims, mu, lo, hi):
    """Sites of arr ([V, ...], lexicographic [t][z][y][x]) with lo <= x_mu < hi."""
    Lx, Ly, Lz, Lt = dims
    a = arr.reshape((Lt, Lz, Ly, Lx) + arr.shape[1:])
    axis = 3 - mu
    sl = [slice(None)] * 4
    sl[axis] = slice(lo, hi)
    sub = a[tuple(sl)]
    return np.ascontiguousarray(sub.reshape((-1,) + arr.shape[1:]))


@pytest.mark.parametrize("mu,cuts", [(0, [0, 2, 6]), (1, [0, 1, 4]), (2, [0, 3, 4])])
def test_spatial_split_adds_up(mu, cuts):
    dims = (6, 4, 4, 3)
    rng = np.random.default_rng(mu)
    V = int(np.prod(dims))
    c = rng.standard_normal((V, 2)) + 1j * rng.standard_normal((V, 2))
    whole = oracle.project(c, dims, MOM)
    total = np.zeros_like(whole)
    for lo, hi in zip(cuts[:-1], cuts[1:]):
        sub_dims = list(dims)
        sub_dims[mu] = hi - lo
        origin = [0, 0, 0]
        origin[mu] = lo
        total += oracle.project(split(c, dims, mu, lo, hi), sub_dims, MOM, global_dims=dims[:3], origin=origin)
    assert np.abs(total - whole).max() < 1e-13 * np.abs(whole).max()


def test_time_split_stacks():
    dims = (4, 4, 2, 6)
    rng = np.random.default_rng(9)
    V = int(np.prod(dims))
    c = rng.standard_normal((V, 1)) + 1j * rng.standard_normal((V, 1))
    whole = oracle.project(c, dims, MOM)
    parts = [oracle.project(split(c, dims, 3, lo, hi), dims[:3] + (hi - lo,), MOM) for lo, hi in [(0, 2), (2, 6)]]
    assert np.abs(np.concatenate(parts, axis=2) - whole).max() < 1e-13 * np.abs(whole).max()


def test_origin_is_a_phase():
    """Shifting the origin multiplies the projection by exp(-i 2pi p.o / G)."""
    dims = (4, 2, 2, 2)
    G = (8, 6, 4)
    o = (3, 5, 1)
    rng = np.random.default_rng(4)
    c = rng.standard_normal((int(np.prod(dims)), 1)) + 0j
    a = oracle.project(c, dims, MOM, global_dims=G, origin=(0, 0, 0))
    b = oracle.project(c, dims, MOM, global_dims=G, origin=o)
    ph = np.exp(-2j * np.pi * np.array([sum(p[m] * o[m] / G[m] for m in range(3)) for p in MOM]))
    assert np.abs(b - a * ph[None, :, None]).max() < 1e-13 * np.abs(a).max()
