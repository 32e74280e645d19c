"""Pin the C oracle against the dense-einsum golden vectors
(tests/golden/golden.npz, written by tests/golden/make_golden.py, which shares
no code with oracle/contract_oracle.c)."""

import numpy as np
import pytest

import oracle

TOL = 1e-13


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def keys(golden, prefix):
    return [k for k in golden.files if k.startswith(prefix)]


def test_meson(golden):
    ks = keys(golden, "meson_")
    assert ks
    for k in ks:
        _, a, b = k.split("_")
        assert rel(oracle.meson_two_point(golden["pi"], golden["pj"], int(a), int(b)), golden[k]) < TOL, k


def test_diquark(golden):
    for k in keys(golden, "diquark_"):
        _, a, b = k.split("_")
        assert rel(oracle.baryon_diquark(golden["pi"], golden["pj"], int(a), int(b)), golden[k]) < TOL, k


def test_baryon_all_types(golden):
    ks = keys(golden, "baryon_")
    assert len(ks) == 24
    for k in ks:
        _, t, a, b, c = k.split("_")
        got = oracle.baryon_two_point(golden["pi"], golden["pj"], golden["pn"], int(t), int(a), int(b), int(c))
        assert rel(got, golden[k]) < TOL, k


def test_general(golden):
    for k in keys(golden, "general_"):
        _, t, a, b = k.split("_")
        got = oracle.baryon_general_two_point(golden["pi"], golden["pj"], golden["pn"], int(t), int(a), int(b),
                                              golden["project_mn"])
        assert rel(got, golden[k]) < TOL, k


def test_sequential(golden):
    ks = keys(golden, "sequential_")
    assert len(ks) == 36
    for k in ks:
        _, t, s, a, b, c = k.split("_")
        outs = oracle.baryon_sequential_two_point(golden["pi"], golden["pj"], golden["pn"], int(t), int(s), int(a),
                                                  int(b), int(c))
        assert rel(outs[int(s)], golden[k]) < TOL, k
