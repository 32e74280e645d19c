"""Write ref_vectors.npz: outputs of the reference's own library on the golden
inputs, so the CPU test suite can check the oracle against the reference
itself (tests/test_oracle_ref_vectors.py).

The reference is HIP-only, so this runs on a GPU machine from the repo root,
after `make -C oracle ref` built oracle/_ref/libcontract_ref.so from the
reference sources:

    python tests/golden/make_ref_vectors.py OUT/ref_vectors.npz

and the file is then copied to tests/golden/.

Inputs are golden.npz's pi, pj, pn (2^3x4 lattice, V = 32).  Only data is
written: the inputs' identity (by name) and the reference's outputs.
"""

import ctypes
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "libcontract_ref.so")


def main(path):
    z = np.load(os.path.join(HERE, "golden.npz"))
    ps = [z["pi"], z["pj"], z["pn"]]
    V = len(ps[0])
    L = ctypes.CDLL(REF, mode=ctypes.RTLD_LOCAL)
    P, U, S, I = ctypes.c_void_p, ctypes.c_ulong, ctypes.c_size_t, ctypes.c_int
    L.init.argtypes = [I]
    L.meson_two_point.argtypes = [P, P, P, U, I, I]
    L.baryon_diquark.argtypes = [P, P, P, S, I, I]
    L.baryon_two_point.argtypes = [P, P, P, P, I, U, I, I, I]
    L.baryon_general_two_point.argtypes = [P, P, P, P, I, S, I, I, P]
    L.baryon_sequential_two_point.argtypes = [P, P, P, I, I, S, I, I, I]
    L.init(0)
    dev = torch.device("cuda:0")
    d = [torch.from_numpy(p).to(dev) for p in ps]
    out = torch.empty(V, dtype=torch.complex128, device=dev)
    res = {}

    def grab(t):
        torch.cuda.synchronize()
        return t.cpu().numpy().copy()

    for gij in range(16):
        for gkl in range(16):
            L.meson_two_point(out.data_ptr(), d[0].data_ptr(), d[1].data_ptr(), V, gij, gkl)
            res[f"meson_{gij}_{gkl}"] = grab(out)
    dq = torch.empty_like(d[0])
    for gij in range(16):
        for gkl in (0, 5, 10, 15):
            L.baryon_diquark(dq.data_ptr(), d[0].data_ptr(), d[1].data_ptr(), V, gij, gkl)
            res[f"diquark_{gij}_{gkl}"] = grab(dq)
    for t in range(6):
        for gij, gkl in [(5, 5), (13, 5), (10, 3)]:
            for gmn in range(16):
                L.baryon_two_point(out.data_ptr(), d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), t, V, gij,
                                   gkl, gmn)
                res[f"baryon_{t}_{gij}_{gkl}_{gmn}"] = grab(out)
        proj = np.ascontiguousarray(z["project_mn"])
        L.baryon_general_two_point(out.data_ptr(), d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), t, V, 5, 5,
                                   proj.ctypes.data)
        res[f"general_{t}_5_5"] = grab(out)
        for s in range(3):
            for gij, gkl, gmn in [(5, 5, 8), (13, 10, 3)]:
                dd = [torch.from_numpy(p).to(dev) for p in ps]
                L.baryon_sequential_two_point(dd[0].data_ptr(), dd[1].data_ptr(), dd[2].data_ptr(), t, s, V, gij, gkl,
                                              gmn)
                res[f"sequential_{t}_{s}_{gij}_{gkl}_{gmn}"] = grab(dd[s])
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez_compressed(path, **res)
    print("wrote", path, len(res), "arrays")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "ref_vectors.npz"))
