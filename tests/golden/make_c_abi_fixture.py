"""Writes tests/golden/c_abi_fixture.bin: the FP64 oracle's results for the
inputs and calls of tests/c/abi_check.c, complex128 arrays one after another
in the order that program checks them.

    python tests/golden/make_c_abi_fixture.py

The inputs are regenerated here with the same splitmix64 construction as the
C program (entry k of propagator `seed` is 2u - 1, u = splitmix64(seed << 32 | k)
/ 2^64 taken as the top 53 bits), so the C program needs no input files."""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

DIMS = (4, 2, 2, 4)
V = 64
MOM = [(0, 0, 0), (1, 0, 0), (0, -1, 1)]
M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    x = x.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def propagator(seed):
    k = np.arange(V * 144 * 2, dtype=np.uint64)
    u = (splitmix64((np.uint64(seed) << np.uint64(32)) | k) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    v = 2.0 * u - 1.0
    return (v[0::2] + 1j * v[1::2]).reshape(V, 4, 4, 3, 3)


def results():
    p = [propagator(s) for s in (1, 2, 3)]
    g4 = np.zeros((4, 4), np.complex128)
    for i in range(4):
        g4[i, (i + 2) % 4] = 1
    out = [
        oracle.baryon_two_point(*p, 0, 5, 5, 8),
        oracle.baryon_two_point(*p, 5, 5, 5, 8),
        oracle.baryon_general_two_point(*p, 1, 5, 5, g4),
        oracle.meson_all_source_two_point(p[0], p[1], 5),
        oracle.meson_two_point(p[0], p[1], 15, 15),
        oracle.baryon_diquark(p[0], p[1], 5, 5),
        oracle.baryon_sequential_two_point(*p, 0, 2, 5, 5, 8)[2],
        oracle.baryon_twopoint_project(*p, [(0, 1.0), (5, 1.0)], 5, 5, list(range(16)), DIMS, MOM),
        oracle.threepoint_project(p[0], p[1], list(range(16)), DIMS, MOM),
    ]
    return np.concatenate([np.ascontiguousarray(o, np.complex128).reshape(-1) for o in out])


if __name__ == "__main__":
    data = results()
    path = os.path.join(HERE, "c_abi_fixture.bin")
    data.tofile(path)
    print(f"wrote {path}: {data.size} complex128")
