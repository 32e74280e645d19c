"""Host-side checks of libcontract.so that need no GPU: the library loads,
exports every function include/*.h declares, and rejects bad arguments the
way the reference does (print + exit(-1), or a silent no-op) before any HIP
call is made."""

import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_headers_declare_expected_abi():
    assert declared_functions("contract.h") == sorted([
        "init", "meson_two_point", "meson_all_source_two_point", "meson_all_sink_two_point", "baryon_diquark",
        "baryon_two_point", "baryon_general_two_point", "baryon_sequential_two_point"])


def test_library_exports_every_declared_symbol():
    import contract_amd
    L = contract_amd.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", contract_amd.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for h in ("contract.h", "contract_ext.h"):
        for name in declared_functions(h):
            assert name in exported, (h, name)
            getattr(L, name)
    assert set(contract_amd.ABI_SYMBOLS) | set(contract_amd.EXT_SYMBOLS) <= exported


def test_enums_match_reference_values():
    from contract_amd.pycontract import BaryonContractType as T, BaryonSequentialType as S
    assert [T.IK_JL_NM, T.IK_JM_NL, T.IL_JK_NM, T.IL_JM_NK, T.IM_JK_NL, T.IM_JL_NK] == list(range(6))
    assert (T.IK_JL_MN, T.IK_JN_ML, T.IL_JK_MN, T.IL_JN_MK, T.IN_JK_ML, T.IN_JL_MK) == tuple(range(6))
    assert (T.AD_BE_CF, T.AD_BF_CE, T.AE_BD_CF, T.AE_BF_CD, T.AF_BD_CE, T.AF_BE_CD) == tuple(range(6))
    assert (S.SEQUENTIAL_I, S.SEQUENTIAL_J, S.SEQUENTIAL_N) == (0, 1, 2)


def test_gamma_matrix_matches_oracle():
    import numpy as np
    import oracle
    from contract_amd.pycontract import gamma_matrix
    for g in range(16):
        assert np.array_equal(gamma_matrix(g), oracle.gamma(g))


def run_abi(snippet):
    code = ("import ctypes, sys; sys.path.insert(0, %r)\n"
            "import contract_amd; L = contract_amd.lib()\n" % ROOT) + snippet + "\nprint('RETURNED')"
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("call,msg", [
    ("L.meson_two_point(None, None, None, 6, 5, 5)", "multiple of TILES_PER_BLOCK"),
    ("L.baryon_two_point(None, None, None, None, 6, 8, 5, 5, 8)", "Invalid contract_type"),
    ("L.baryon_two_point(None, None, None, None, 0, 8, 5, 5, 16)", "Invalid gamma_mn"),
    ("L.baryon_two_point(None, None, None, None, 0, 10, 5, 5, 8)", "multiple of TILES_PER_BLOCK"),
    ("L.baryon_two_point(None, None, None, None, 0, 8, 17, 5, 8)", "Invalid gamma_ij"),
    ("L.meson_two_point(None, None, None, 8, 5, -1)", "Invalid gamma_kl"),
    ("L.baryon_sequential_two_point(None, None, None, 9, 0, 8, 5, 5, 8)", "Invalid contract_type"),
    ("L.baryon_diquark(None, None, None, 7, 5, 5)", "multiple of TILES_PER_BLOCK"),
])
def test_invalid_arguments_exit(call, msg):
    r = run_abi(call)
    assert r.returncode == 255, (r.returncode, r.stdout, r.stderr)
    assert msg in r.stderr
    assert "RETURNED" not in r.stdout


@pytest.mark.parametrize("call", [
    "L.baryon_sequential_two_point(None, None, None, 0, 3, 8, 5, 5, 8)",   # lib/contract.cpp:60
    "L.baryon_diquark(None, None, None, 8, 16, 5)",                        # lib/baryon/diquark.cu:40
])
def test_silent_noops(call):
    r = run_abi(call)
    assert r.returncode == 0 and "RETURNED" in r.stdout, (r.stdout, r.stderr)


def test_extension_rejects_bad_arguments_with_codes():
    import ctypes
    import numpy as np
    import contract_amd
    L = contract_amd.lib()
    vp = ctypes.c_void_p
    dims = np.array([4, 4, 4, 8], np.int32)
    mom = np.zeros((1, 3), np.int32)
    types = np.array([0], np.int32)
    w = np.array([1.0], np.complex128)
    proj = np.eye(4, dtype=np.complex128)

    def call(**kw):
        a = dict(out=1, pi=1, pj=1, pn=1, prec=0, types=types.ctypes.data, w=w.ctypes.data, nterm=1, gij=5, gkl=5,
                 proj=proj.ctypes.data, nproj=1, mom=mom.ctypes.data, nmom=1, dims=dims.ctypes.data, order=0)
        a.update(kw)
        return L.contract_baryon_twopoint_project(a["out"], a["pi"], a["pj"], a["pn"], a["prec"], a["types"],
                                                  a["w"], a["nterm"], a["gij"], a["gkl"], a["proj"], a["nproj"],
                                                  a["mom"], a["nmom"], a["dims"], a["order"], None, 0, None)

    assert call(nmom=0) != 0 and "nmom" in contract_amd.last_error()
    assert call(prec=3) != 0 and "precision" in contract_amd.last_error()
    assert call(nterm=7) != 0
    bad = np.array([9], np.int32)
    assert call(types=bad.ctypes.data) != 0 and "BaryonContractType" in contract_amd.last_error()
    assert call(gij=16) != 0
    assert call(order=2) != 0 and "site_order" in contract_amd.last_error()
    odd = np.array([3, 4, 4, 8], np.int32)
    assert call(dims=odd.ctypes.data, order=1) != 0 and "even" in contract_amd.last_error()
    assert call(out=0) != 0
    assert L.contract_project_workspace_bytes(dims.ctypes.data, 0, 27) > 0
    assert L.contract_project_workspace_bytes(dims.ctypes.data, 0, 0) == 0


def test_baryon_batch_validation_and_workspace():
    import ctypes
    import numpy as np
    import contract_amd
    from contract_amd import pycontract as pc
    L = contract_amd.lib()
    vp = ctypes.c_void_p
    mom = np.zeros((1, 3), np.int32)
    types = np.array([0], np.int32)
    w = np.array([1.0], np.complex128)
    proj = np.eye(4, dtype=np.complex128)

    def call(arr, n):
        return L.contract_baryon_twopoint_project_batch(arr, n, 0, types.ctypes.data, w.ctypes.data, 1, 5, 5,
                                                        proj.ctypes.data, 1, mom.ctypes.data, 1, None, 0, None)

    assert call(None, 1) != 0 and "nreq" in contract_amd.last_error()
    arr = (pc.BaryonRequest * 2)()
    for r in arr:
        r.propag_i = r.propag_j = r.propag_n = r.out = 1
        r.lattice.dims[:] = [4, 4, 4, 8]
    arr[1].propag_n = None
    assert call(ctypes.cast(arr, vp), 2) != 0 and "request 1" in contract_amd.last_error()
    arr[1].propag_n = 1
    arr[1].lattice.site_order = 1
    arr[1].lattice.dims[0] = 3
    assert call(ctypes.cast(arr, vp), 2) != 0 and "even" in contract_amd.last_error()
    # the batch's workspace is the sum of its requests' own
    a = (pc.BaryonRequest * 2)()
    a[0].lattice.dims[:] = [4, 4, 4, 8]
    a[1].lattice.dims[:] = [6, 4, 2, 6]
    a[1].lattice.site_order = 1
    single = [L.contract_project_workspace_bytes(np.array(d, np.int32).ctypes.data, o, 27)
              for d, o in (([4, 4, 4, 8], 0), ([6, 4, 2, 6], 1))]
    assert L.contract_baryon_batch_workspace_bytes(ctypes.cast(a, vp), 2, 27) == sum(single) > 0
    assert L.contract_baryon_batch_workspace_bytes(ctypes.cast(a, vp), 2, 0) == 0


def test_scaled_fp32_form_size_and_validation():
    import contract_amd
    L = contract_amd.lib()
    assert L.contract_scaled_fp32_bytes(1000) == 1153 * 1000
    assert L.contract_propagator_to_scaled_fp32(None, None, 10, None) != 0
    assert "NULL" in contract_amd.last_error()


def test_meson_grid_rejects_bad_arguments_with_codes():
    import ctypes
    import numpy as np
    import contract_amd
    from contract_amd import pycontract as pc
    L = contract_amd.lib()
    mom = np.zeros((1, 3), np.int32)

    def call(dims=(4, 4, 4, 8), order=0, nreq=1, prec=0, nmom=1, tuning=None, ptrs=(1, 1, 1)):
        arr = (pc.MesonRequest * max(nreq, 1))()
        for r in range(max(nreq, 1)):
            arr[r].propag_i, arr[r].propag_j, arr[r].out = ptrs
            for mu in range(4):
                arr[r].dims[mu] = dims[mu]
            arr[r].site_order = order
        t = pc.MesonGridTuning(*tuning) if tuning else None
        return L.contract_meson_grid_project(arr, nreq, prec, mom.ctypes.data, nmom,
                                             ctypes.byref(t) if t else None, None, 0, None)

    assert call(nreq=0) != 0 and "nreq" in contract_amd.last_error()
    assert call(prec=3) != 0 and "precision" in contract_amd.last_error()
    assert call(nmom=0) != 0 and "nmom" in contract_amd.last_error()
    assert call(tuning=(0, 6, 0)) != 0 and "tile_sites" in contract_amd.last_error()
    assert call(tuning=(0, 0, 9)) != 0 and "blocks_per_cu" in contract_amd.last_error()
    assert call(tuning=(-5, 0, 0)) != 0 and "sites_per_item" in contract_amd.last_error()
    assert call(dims=(3, 4, 4, 8), order=1) != 0 and "even" in contract_amd.last_error()
    assert call(dims=(200, 200, 4, 8)) != 0 and "exceeds" in contract_amd.last_error()
    assert call(ptrs=(1, 0, 1)) != 0 and "NULL" in contract_amd.last_error()
    reqs = [(None, None, (16, 16, 16, 32)), (None, None, (24, 24, 24, 48), 1)]
    # 16^3 x 32: 32 slices x 4 items; 24^3 x 48 even-odd: 96 slices x 7 items (1024 sites per item)
    assert pc.meson_grid_workspace_bytes(reqs, 7) == (32 * 4 + 96 * 7) * 16 * 2 * 256 * 8
    assert pc.meson_grid_workspace_bytes(reqs, 7, (4096, 8, 1)) == (32 * 1 + 96 * 2) * 16 * 2 * 256 * 8
    assert pc.meson_grid_workspace_bytes(reqs, 0) == 0
