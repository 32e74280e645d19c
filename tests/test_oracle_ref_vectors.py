"""Pin the oracle against outputs of the reference's own HIP library
(tests/golden/ref_vectors.npz, written on an MI355X by
tests/golden/make_ref_vectors.py from a build of the reference's sources) on
the golden inputs: 256 meson, 64 diquark, 288 baryon, 6 general and 36
sequential vectors."""

import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "golden", "ref_vectors.npz")
TOL = 1e-12


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(PATH):
        pytest.skip("ref_vectors.npz not generated yet")
    return np.load(PATH)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_every_reference_vector(ref, golden):
    pi, pj, pn = golden["pi"], golden["pj"], golden["pn"]
    n = 0
    for k in ref.files:
        parts = k.split("_")
        kind, args = parts[0], [int(v) for v in parts[1:]]
        if kind == "meson":
            got = oracle.meson_two_point(pi, pj, *args)
        elif kind == "diquark":
            got = oracle.baryon_diquark(pi, pj, *args)
        elif kind == "baryon":
            got = oracle.baryon_two_point(pi, pj, pn, *args)
        elif kind == "general":
            got = oracle.baryon_general_two_point(pi, pj, pn, *args, golden["project_mn"])
        elif kind == "sequential":
            got = oracle.baryon_sequential_two_point(pi, pj, pn, *args)[args[1]]
        else:
            raise AssertionError(k)
        assert rel(got, ref[k]) < TOL, k
        n += 1
    assert n == 256 + 64 + 288 + 6 + 36
