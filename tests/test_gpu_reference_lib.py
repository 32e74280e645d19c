"""Cross-check against the reference's own HIP library, built from its sources
into oracle/_ref by `make -C oracle ref`.  On identical device inputs the
reference and this library must agree to FP64 rounding, and both must agree
with the CPU oracle.  Skipped when oracle/_ref was not built."""

import ctypes
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def ref(cuda):
    path = oracle.ref_library_path()
    if not os.path.exists(path):
        pytest.skip("oracle/_ref/libcontract_ref.so not built")
    L = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
    P, U, S, I = ctypes.c_void_p, ctypes.c_ulong, ctypes.c_size_t, ctypes.c_int
    L.init.argtypes = [I]
    L.meson_two_point.argtypes = [P, P, P, U, I, I]
    L.meson_all_source_two_point.argtypes = [ctypes.POINTER(P), P, P, U, I]
    L.meson_all_sink_two_point.argtypes = [ctypes.POINTER(P), P, P, U, I]
    L.baryon_diquark.argtypes = [P, P, P, S, I, I]
    L.baryon_two_point.argtypes = [P, P, P, P, I, U, I, I, I]
    L.baryon_general_two_point.argtypes = [P, P, P, P, I, S, I, I, P]
    L.baryon_sequential_two_point.argtypes = [P, P, P, I, I, S, I, I, I]
    L.init(0)
    return L


@pytest.fixture(scope="module")
def data(cuda):
    import torch
    dims = (4, 4, 2, 4)
    ps = [oracle.synthetic_propagator(dims, s, 0.25) for s in (41, 42, 43)]
    return torch, cuda, ps


def test_baryon_reference_vs_library(ref, data):
    torch, cuda, ps = data
    from contract_amd import pycontract as pc
    d = [torch.from_numpy(p).to(cuda) for p in ps]
    V = len(ps[0])
    out = torch.empty(V, dtype=torch.complex128, device=cuda)
    for t in range(6):
        for gij, gkl, gmn in [(5, 5, 8), (13, 5, 0), (10, 3, 7)]:
            ref.baryon_two_point(out.data_ptr(), d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), t, V, gij, gkl,
                                 gmn)
            torch.cuda.synchronize()
            r = out.cpu().numpy()
            mine = pc.baryonTwoPoint(*d, t, gij, gkl, gmn).cpu().numpy()
            want = oracle.baryon_two_point(*ps, t, gij, gkl, gmn)
            assert rel(r, want) < TOL, ("reference vs oracle", t, gij, gkl, gmn)
            assert rel(mine, r) < TOL, ("library vs reference", t, gij, gkl, gmn)


def test_general_reference_vs_library(ref, data):
    torch, cuda, ps = data
    from contract_amd import pycontract as pc
    d = [torch.from_numpy(p).to(cuda) for p in ps]
    V = len(ps[0])
    out = torch.empty(V, dtype=torch.complex128, device=cuda)
    proj = np.ascontiguousarray(0.5 * (np.eye(4) - oracle.gamma(8)) + 0.3j * oracle.gamma(3))
    for t in range(6):
        ref.baryon_general_two_point(out.data_ptr(), d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), t, V, 5, 5,
                                     proj.ctypes.data)
        torch.cuda.synchronize()
        r = out.cpu().numpy()
        assert rel(r, oracle.baryon_general_two_point(*ps, t, 5, 5, proj)) < TOL
        assert rel(pc.baryonTwoPoint(*d, t, 5, 5, proj).cpu().numpy(), r) < TOL


def test_meson_and_diquark_reference_vs_library(ref, data):
    torch, cuda, ps = data
    from contract_amd import pycontract as pc
    d = [torch.from_numpy(p).to(cuda) for p in ps]
    V = len(ps[0])
    out = torch.empty(V, dtype=torch.complex128, device=cuda)
    dq = torch.empty_like(d[0])
    for gij, gkl in [(5, 5), (0, 15), (1, 8), (13, 10)]:
        ref.meson_two_point(out.data_ptr(), d[0].data_ptr(), d[1].data_ptr(), V, gij, gkl)
        torch.cuda.synchronize()
        r = out.cpu().numpy()
        assert rel(r, oracle.meson_two_point(ps[0], ps[1], gij, gkl)) < TOL
        assert rel(pc.mesonTwoPoint(d[0], d[1], gij, gkl).cpu().numpy(), r) < TOL
        ref.baryon_diquark(dq.data_ptr(), d[0].data_ptr(), d[1].data_ptr(), V, gij, gkl)
        torch.cuda.synchronize()
        r = dq.cpu().numpy()
        assert rel(r, oracle.baryon_diquark(ps[0], ps[1], gij, gkl)) < TOL
        assert rel(pc.baryonDiquark(d[0], d[1], gij, gkl).cpu().numpy(), r) < TOL
    outs = [torch.empty(V, dtype=torch.complex128, device=cuda) for _ in range(16)]
    arr = (ctypes.c_void_p * 16)(*[o.data_ptr() for o in outs])
    ref.meson_all_source_two_point(arr, d[0].data_ptr(), d[1].data_ptr(), V, 5)
    torch.cuda.synchronize()
    r = np.stack([o.cpu().numpy() for o in outs])
    assert rel(r, oracle.meson_all_source_two_point(ps[0], ps[1], 5)) < TOL
    ref.meson_all_sink_two_point(arr, d[0].data_ptr(), d[1].data_ptr(), V, 9)
    torch.cuda.synchronize()
    r = np.stack([o.cpu().numpy() for o in outs])
    assert rel(r, oracle.meson_all_sink_two_point(ps[0], ps[1], 9)) < TOL


@pytest.mark.parametrize("s", range(3))
def test_sequential_reference_vs_library(ref, data, s):
    torch, cuda, ps = data
    from contract_amd import pycontract as pc
    V = len(ps[0])
    for t in range(6):
        d_ref = [torch.from_numpy(p).to(cuda) for p in ps]
        ref.baryon_sequential_two_point(d_ref[0].data_ptr(), d_ref[1].data_ptr(), d_ref[2].data_ptr(), t, s, V, 5, 5,
                                        8)
        torch.cuda.synchronize()
        r = d_ref[s].cpu().numpy()
        want = oracle.baryon_sequential_two_point(*ps, t, s, 5, 5, 8)[s]
        assert rel(r, want) < TOL, ("reference vs oracle", t, s)
        d = [torch.from_numpy(p).to(cuda) for p in ps]
        pc.baryonSequentialTwoPoint(*d, t, s, 5, 5, 8)
        assert rel(d[s].cpu().numpy(), r) < TOL, ("library vs reference", t, s)
