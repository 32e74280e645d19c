"""ctypes loader for libcontract.so (the gfx950 build of the contraction ABI).

The library is built in-tree by ``contract_amd.build`` (or
``__graft_entry__.build()``).  There is no fallback: if the shared object is
missing, importing an entry point raises instead of computing anything on the
CPU.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CONTRACT_AMD_LIB") or os.path.join(HERE, "lib", "libcontract.so")

_lib = None

# Every symbol include/contract.h and include/contract_ext.h declare.
ABI_SYMBOLS = (
    "init",
    "meson_two_point",
    "meson_all_source_two_point",
    "meson_all_sink_two_point",
    "baryon_diquark",
    "baryon_two_point",
    "baryon_general_two_point",
    "baryon_sequential_two_point",
)
EXT_SYMBOLS = (
    "contract_last_error",
    "contract_gamma_matrix",
    "contract_project_workspace_bytes",
    "contract_baryon_twopoint_project",
    "contract_baryon_twopoint_project_sub",
    "contract_baryon_batch_workspace_bytes",
    "contract_scaled_fp32_bytes",
    "contract_propagator_to_scaled_fp32",
    "contract_baryon_twopoint_project_batch",
    "contract_set_kernel_timer",
    "contract_set_stream",
    "contract_meson_grid_workspace_bytes",
    "contract_meson_grid_project",
    "contract_set_baryon_plan",
    "contract_get_baryon_plan",
    "contract_threepoint_workspace_bytes",
    "contract_threepoint_project",
    "contract_noalloc_scratch_bytes",
    "contract_baryon_twopoint_project_noalloc",
    "contract_baryon_sequential",
    "contract_baryon_diquark",
)


class LibraryMissing(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    """The loaded libcontract.so with argument types set."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(
            f"{LIB_PATH} is not built; run `python -m contract_amd.build` (there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    P, U, S, I = ctypes.c_void_p, ctypes.c_ulong, ctypes.c_size_t, ctypes.c_int
    PP = ctypes.POINTER(ctypes.c_void_p)
    L.init.argtypes = [I]
    L.init.restype = None
    L.meson_two_point.argtypes = [P, P, P, U, I, I]
    L.meson_two_point.restype = None
    L.meson_all_source_two_point.argtypes = [PP, P, P, U, I]
    L.meson_all_source_two_point.restype = None
    L.meson_all_sink_two_point.argtypes = [PP, P, P, U, I]
    L.meson_all_sink_two_point.restype = None
    L.baryon_diquark.argtypes = [P, P, P, S, I, I]
    L.baryon_diquark.restype = None
    L.baryon_two_point.argtypes = [P, P, P, P, I, U, I, I, I]
    L.baryon_two_point.restype = None
    L.baryon_general_two_point.argtypes = [P, P, P, P, I, S, I, I, P]
    L.baryon_general_two_point.restype = None
    L.baryon_sequential_two_point.argtypes = [P, P, P, I, I, S, I, I, I]
    L.baryon_sequential_two_point.restype = None
    L.contract_last_error.argtypes = []
    L.contract_last_error.restype = ctypes.c_char_p
    L.contract_gamma_matrix.argtypes = [I, P]
    L.contract_gamma_matrix.restype = I
    L.contract_project_workspace_bytes.argtypes = [P, I, I]
    L.contract_project_workspace_bytes.restype = S
    L.contract_baryon_twopoint_project.argtypes = [P, P, P, P, I, P, P, I, I, I, P, I, P, I, P, I, P, S, P]
    L.contract_baryon_twopoint_project.restype = I
    L.contract_baryon_twopoint_project_sub.argtypes = [P, P, P, P, I, P, P, I, I, I, P, I, P, I, P, P, S, P]
    L.contract_baryon_twopoint_project_sub.restype = I
    L.contract_scaled_fp32_bytes.argtypes = [S]
    L.contract_scaled_fp32_bytes.restype = S
    L.contract_propagator_to_scaled_fp32.argtypes = [P, P, S, P]
    L.contract_propagator_to_scaled_fp32.restype = I
    L.contract_baryon_batch_workspace_bytes.argtypes = [P, I, I]
    L.contract_baryon_batch_workspace_bytes.restype = S
    L.contract_baryon_twopoint_project_batch.argtypes = [P, I, I, P, P, I, I, I, P, I, P, I, P, S, P]
    L.contract_baryon_twopoint_project_batch.restype = I
    L.contract_set_kernel_timer.argtypes = [P, P]
    L.contract_set_kernel_timer.restype = I
    L.contract_set_stream.argtypes = [P]
    L.contract_set_stream.restype = I
    L.contract_meson_grid_workspace_bytes.argtypes = [P, I, I, P]
    L.contract_meson_grid_workspace_bytes.restype = S
    L.contract_meson_grid_project.argtypes = [P, I, I, P, I, P, P, S, P]
    L.contract_meson_grid_project.restype = I
    L.contract_set_baryon_plan.argtypes = [P]
    L.contract_set_baryon_plan.restype = I
    L.contract_get_baryon_plan.argtypes = [P]
    L.contract_get_baryon_plan.restype = I
    L.contract_threepoint_workspace_bytes.argtypes = [P, I, I]
    L.contract_threepoint_workspace_bytes.restype = S
    L.contract_threepoint_project.argtypes = [P, I, I, P, I, P, I, P, S, P]
    L.contract_threepoint_project.restype = I
    L.contract_noalloc_scratch_bytes.argtypes = []
    L.contract_noalloc_scratch_bytes.restype = S
    L.contract_baryon_twopoint_project_noalloc.argtypes = [P, P, P, P, I, P, P, I, I, I, P, I, P, I, P, P]
    L.contract_baryon_twopoint_project_noalloc.restype = I
    L.contract_baryon_sequential.argtypes = [P, P, P, I, I, I, S, I, I, I, P]
    L.contract_baryon_sequential.restype = I
    L.contract_baryon_diquark.argtypes = [P, P, P, I, S, I, I, P]
    L.contract_baryon_diquark.restype = I
    _lib = L
    return L


def last_error() -> str:
    return lib().contract_last_error().decode()
