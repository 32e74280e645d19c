"""Python face of libcontract.so, shaped like the reference's plugin binding.

The reference is consumed from Python through a module that
``python -m pyquda_plugins -i contract.h -l contract`` generates from
include/contract.h (README.md:13); PyQUDA scripts call it as
``pycontract.baryonTwoPoint(...)`` (baryon_2pt.py:92,108).  This module offers
the same surface over the gfx950 library:

* the C entry points under their C names with the C argument order
  (``meson_two_point(correl, propag_i, propag_j, volume, gamma_ij, gamma_kl)``
  ...), accepting device pointers as ints or torch tensors;
* the ``BaryonContractType`` / ``BaryonSequentialType`` enums with every
  alias of include/contract_enum.h;
* tensor-level helpers (``mesonTwoPoint``, ``baryonTwoPoint``, ...) that take
  torch complex128 propagators ``[V, 4, 4, 3, 3]`` on the GPU and return the
  correlator, plus ``baryonTwoPointProject`` for the fused timeslice/momentum
  extension.

Launches go to the null stream, like the reference's; synchronise before
reading results on the host.
"""

from __future__ import annotations

import ctypes
import enum
from typing import Sequence

import numpy as np

from ._lib import last_error, lib


class BaryonContractType(enum.IntEnum):
    IK_JL_NM = 0
    IK_JM_NL = 1
    IL_JK_NM = 2
    IL_JM_NK = 3
    IM_JK_NL = 4
    IM_JL_NK = 5
    AD_BE_CF = 0
    AD_BF_CE = 1
    AE_BD_CF = 2
    AE_BF_CD = 3
    AF_BD_CE = 4
    AF_BE_CD = 5
    IK_JL_MN = 0
    IK_JN_ML = 1
    IL_JK_MN = 2
    IL_JN_MK = 3
    IN_JK_ML = 4
    IN_JL_MK = 5


class BaryonSequentialType(enum.IntEnum):
    SEQUENTIAL_I = 0
    SEQUENTIAL_J = 1
    SEQUENTIAL_N = 2


CONTRACT_FP64 = 0
CONTRACT_FP32 = 1
CONTRACT_FP32_SCALED = 2
LEXICOGRAPHIC = 0
EVEN_ODD = 1


def _ptr(x) -> int:
    """Device pointer of a torch tensor, or an int passed through."""
    if x is None:
        return 0
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    raise TypeError(f"expected a device pointer or tensor, got {type(x)}")


def _nbytes(x) -> int:
    if isinstance(x, ScaledPropagator):
        x = x.data
    return x.numel() * x.element_size()


def _spans(*args):
    """Raw-ABI argument check over (x, need_bytes, name) triples: a tensor
    passed where the C entry point takes a device pointer must be contiguous
    and hold at least `need` bytes (checked for every argument first), and all
    such tensors must be on one GPU (an int pointer is the caller's
    responsibility, as in C)."""
    tensors = []
    for x, need, what in args:
        if x is None or isinstance(x, int):
            continue
        t = x.data if isinstance(x, ScaledPropagator) else x
        if not hasattr(t, "is_contiguous"):
            continue
        if not t.is_contiguous():
            raise ValueError(f"{what} must be contiguous")
        if _nbytes(t) < need:
            raise ValueError(f"{what} holds {_nbytes(t)} bytes, the call reads or writes {need}")
        tensors.append(t)
    _same_device(tensors, "the tensor arguments")


def _same_device(tensors, what: str):
    """Every tensor is on one GPU; checked after the size checks so that host
    tests can reach those without a device."""
    devs = []
    for t in tensors:
        if t is None:
            continue
        d = t.data.device if isinstance(t, ScaledPropagator) else t.device
        if d.type != "cuda":
            raise ValueError(f"{what} must be GPU tensors, got device {d}")
        devs.append(d)
    if len(set(devs)) > 1:
        raise ValueError(f"{what} must share one GPU, got {sorted(set(str(d) for d in devs))}")


def _check_out(out, shape, what: str = "out"):
    import torch
    if out.dtype != torch.complex128:
        raise ValueError(f"{what} must be complex128, got {out.dtype}")
    if tuple(out.shape) != tuple(shape):
        raise ValueError(f"{what} has shape {tuple(out.shape)}, expected {tuple(shape)}")
    if not out.is_contiguous():
        raise ValueError(f"{what} must be contiguous")


def _check_workspace(ws, need: int = 0):
    if ws is None:
        return
    if not ws.is_contiguous():
        raise ValueError("workspace must be contiguous")
    if need and _nbytes(ws) < need:
        raise ValueError(f"workspace of {_nbytes(ws)} bytes is smaller than the {need} needed")


# ---------------------------------------------------------------------------
# C entry points (include/contract.h)
# ---------------------------------------------------------------------------

def init(device: int) -> None:
    lib().init(int(device))


SITE_BYTES_FP64 = 2304  # complex128 [4][4][3][3]


def _props_args(volume, **props):
    return [(x, SITE_BYTES_FP64 * int(volume), n) for n, x in props.items()]


def meson_two_point(correl, propag_i, propag_j, volume: int, gamma_ij: int, gamma_kl: int) -> None:
    _spans((correl, 16 * int(volume), "correl"), *_props_args(volume, propag_i=propag_i, propag_j=propag_j))
    lib().meson_two_point(_ptr(correl), _ptr(propag_i), _ptr(propag_j), int(volume), int(gamma_ij), int(gamma_kl))


def _ptr_array(correl) -> ctypes.Array:
    ptrs = [_ptr(c) for c in correl]
    if len(ptrs) != 16:
        raise ValueError("meson_all_* need 16 output pointers")
    return (ctypes.c_void_p * 16)(*ptrs)


def _span_all(correl, propag_i, propag_j, volume):
    _spans(*[(c, 16 * int(volume), f"correl[{g}]") for g, c in enumerate(correl)],
           *_props_args(volume, propag_i=propag_i, propag_j=propag_j))


def meson_all_source_two_point(correl, propag_i, propag_j, volume: int, gamma_ij: int) -> None:
    _span_all(correl, propag_i, propag_j, volume)
    lib().meson_all_source_two_point(_ptr_array(correl), _ptr(propag_i), _ptr(propag_j), int(volume), int(gamma_ij))


def meson_all_sink_two_point(correl, propag_i, propag_j, volume: int, gamma_kl: int) -> None:
    _span_all(correl, propag_i, propag_j, volume)
    lib().meson_all_sink_two_point(_ptr_array(correl), _ptr(propag_i), _ptr(propag_j), int(volume), int(gamma_kl))


def baryon_diquark(diquark, propag_i, propag_j, volume: int, gamma_ij: int, gamma_kl: int) -> None:
    _spans(*_props_args(volume, diquark=diquark, propag_i=propag_i, propag_j=propag_j))
    lib().baryon_diquark(_ptr(diquark), _ptr(propag_i), _ptr(propag_j), int(volume), int(gamma_ij), int(gamma_kl))


def baryon_two_point(correl, propag_i, propag_j, propag_n, contract_type, volume: int, gamma_ij: int,
                     gamma_kl: int, gamma_mn: int) -> None:
    _spans((correl, 16 * int(volume), "correl"),
           *_props_args(volume, propag_i=propag_i, propag_j=propag_j, propag_n=propag_n))
    lib().baryon_two_point(_ptr(correl), _ptr(propag_i), _ptr(propag_j), _ptr(propag_n), int(contract_type),
                           int(volume), int(gamma_ij), int(gamma_kl), int(gamma_mn))


def baryon_general_two_point(correl, propag_i, propag_j, propag_n, contract_type, volume: int, gamma_ij: int,
                             gamma_kl: int, project_mn) -> None:
    proj = np.ascontiguousarray(np.asarray(project_mn, dtype=np.complex128).reshape(16))
    _spans((correl, 16 * int(volume), "correl"),
           *_props_args(volume, propag_i=propag_i, propag_j=propag_j, propag_n=propag_n))
    lib().baryon_general_two_point(_ptr(correl), _ptr(propag_i), _ptr(propag_j), _ptr(propag_n),
                                   int(contract_type), int(volume), int(gamma_ij), int(gamma_kl),
                                   proj.ctypes.data_as(ctypes.c_void_p))


def baryon_sequential_two_point(propag_i, propag_j, propag_n, contract_type, sequential_type, volume: int,
                                gamma_ij: int, gamma_kl: int, gamma_mn: int) -> None:
    _spans(*_props_args(volume, propag_i=propag_i, propag_j=propag_j, propag_n=propag_n))
    lib().baryon_sequential_two_point(_ptr(propag_i), _ptr(propag_j), _ptr(propag_n), int(contract_type),
                                      int(sequential_type), int(volume), int(gamma_ij), int(gamma_kl),
                                      int(gamma_mn))


# ---------------------------------------------------------------------------
# tensor-level helpers
# ---------------------------------------------------------------------------

def gamma_matrix(g: int) -> np.ndarray:
    """Dense DeGrand-Rossi G(g) as computed by the library."""
    out = np.zeros((4, 4), np.complex128)
    if lib().contract_gamma_matrix(int(g), out.ctypes.data_as(ctypes.c_void_p)) != 0:
        raise ValueError(last_error())
    return out


class ScaledPropagator:
    """A propagator in the FP32 site-scaled form (include/contract_ext.h):
    `data` is a uint8 device tensor of V * 1153 bytes, complex64 P(x) 2^-e(x)
    followed by the int8 site exponents e(x)."""

    dtype = "fp32_scaled"  # storage form tag (not a torch dtype)

    def __init__(self, data, volume: int):
        self.data, self.volume = data, int(volume)
        self.device = data.device

    def data_ptr(self) -> int:
        return self.data.data_ptr()

    def entries(self):
        """The stored complex64 [V, 4, 4, 3, 3] (scaled) entries, a view."""
        return self.data[: self.volume * 1152].view(dtype=__import__("torch").complex64).view(-1, 4, 4, 3, 3)

    def exponents(self):
        """The int8 [V] site exponents, a view."""
        return self.data[self.volume * 1152:].view(dtype=__import__("torch").int8)

    def to_complex128(self):
        """The represented propagator, entries x 2^e(x), as complex128 [V, 4, 4, 3, 3]."""
        import torch
        scale = torch.ldexp(torch.ones(self.volume, dtype=torch.float64, device=self.device),
                            self.exponents().to(torch.int32))
        return (self.entries().to(torch.complex128) * scale.view(-1, 1, 1, 1, 1)).contiguous()

    def clone(self):
        return ScaledPropagator(self.data.clone(), self.volume)


def scaled_fp32_bytes(volume: int) -> int:
    return int(lib().contract_scaled_fp32_bytes(int(volume)))


def toScaledFP32(propag, out=None, stream=None) -> ScaledPropagator:
    """complex128 [V, 4, 4, 3, 3] device propagator -> ScaledPropagator."""
    import torch
    if propag.dtype != torch.complex128:
        raise TypeError("toScaledFP32 converts complex128 propagators")
    V = _volume(propag)
    if out is None:
        out = torch.empty(scaled_fp32_bytes(V), dtype=torch.uint8, device=propag.device)
    _spans((out, V * 1153, "out"), (propag, V * SITE_BYTES_FP64, "propag"))
    rc = lib().contract_propagator_to_scaled_fp32(_ptr(out), _ptr(propag), V, 0 if stream is None else int(stream))
    if rc != 0:
        raise RuntimeError(f"contract_propagator_to_scaled_fp32 failed ({rc}): {last_error()}")
    return ScaledPropagator(out, V)


def _precision(props) -> int:
    """Precision flag of a set of propagators that must share one storage form."""
    import torch
    kinds = {"scaled" if isinstance(p, ScaledPropagator) else str(p.dtype) for p in props}
    if len(kinds) != 1:
        raise TypeError(f"propagators of one call must share a storage form, got {sorted(kinds)}")
    kind = kinds.pop()
    if kind == "scaled":
        return CONTRACT_FP32_SCALED
    for p in props:
        _volume(p)
    return CONTRACT_FP32 if kind == str(torch.complex64) else CONTRACT_FP64


def _volume(p) -> int:
    if isinstance(p, ScaledPropagator):
        if not p.data.is_contiguous():
            raise ValueError("a ScaledPropagator's data must be contiguous")
        if _nbytes(p.data) < p.volume * 1153:
            raise ValueError(f"ScaledPropagator of {p.volume} sites holds {_nbytes(p.data)} bytes, "
                             f"the form needs {p.volume * 1153}")
        return p.volume
    if p.dtype.__str__() not in ("torch.complex128", "torch.complex64"):
        raise TypeError(f"propagators must be complex128/complex64 tensors, got {p.dtype}")
    if not p.is_contiguous():
        raise ValueError("propagators must be contiguous [V, 4, 4, 3, 3]")
    if p.numel() % 144:
        raise ValueError("propagator size is not a multiple of 144 complex entries")
    return p.numel() // 144


def _fp64_volume(*props) -> int:
    """Site count of the drop-in helpers' complex128 propagators, which must
    all hold the same number of sites on one GPU."""
    import torch
    for p in props:
        if isinstance(p, ScaledPropagator) or p.dtype != torch.complex128:
            raise TypeError("the drop-in entry points take complex128 propagators")
    vols = [_volume(p) for p in props]
    if len(set(vols)) != 1:
        raise ValueError(f"propagators hold different site counts {vols}")
    _same_device(props, "propagators")
    return vols[0]


def _volume_of_dims(dims) -> int:
    d = [int(v) for v in dims]
    if len(d) != 4:
        raise ValueError(f"dims must be (Lx, Ly, Lz, Lt), got {dims}")
    return d[0] * d[1] * d[2] * d[3]


def _check_lattice_props(props, dims, what: str):
    """Every propagator of a projected call holds exactly prod(dims) sites."""
    V = _volume_of_dims(dims)
    for k, p in enumerate(props):
        if _volume(p) != V:
            raise ValueError(f"{what}: propagator {k} holds {_volume(p)} sites, dims {tuple(dims)} give {V}")
    return V


def mesonTwoPoint(propag_i, propag_j, gamma_ij: int, gamma_kl: int):
    import torch
    V = _fp64_volume(propag_i, propag_j)
    out = torch.empty(V, dtype=torch.complex128, device=propag_i.device)
    meson_two_point(out, propag_i, propag_j, V, gamma_ij, gamma_kl)
    return out


def mesonAllSourceTwoPoint(propag_i, propag_j, gamma_ij: int):
    import torch
    V = _fp64_volume(propag_i, propag_j)
    out = torch.empty((16, V), dtype=torch.complex128, device=propag_i.device)
    meson_all_source_two_point([out[g] for g in range(16)], propag_i, propag_j, V, gamma_ij)
    return out


def mesonAllSinkTwoPoint(propag_i, propag_j, gamma_kl: int):
    import torch
    V = _fp64_volume(propag_i, propag_j)
    out = torch.empty((16, V), dtype=torch.complex128, device=propag_i.device)
    meson_all_sink_two_point([out[g] for g in range(16)], propag_i, propag_j, V, gamma_kl)
    return out


def baryonDiquark(propag_i, propag_j, gamma_ij: int, gamma_kl: int):
    import torch
    V = _fp64_volume(propag_i, propag_j)
    out = torch.empty((V, 4, 4, 3, 3), dtype=torch.complex128, device=propag_i.device)
    baryon_diquark(out, propag_i, propag_j, V, gamma_ij, gamma_kl)
    return out


def baryonTwoPoint(propag_i, propag_j, propag_n, contract_type, gamma_ij: int, gamma_kl: int, gamma_mn):
    """Per-site baryon correlator.  ``gamma_mn`` is a gamma index (sparse
    path) or a 4x4 projector matrix, row-major [m][n] (general path)."""
    import torch
    V = _fp64_volume(propag_i, propag_j, propag_n)
    out = torch.empty(V, dtype=torch.complex128, device=propag_i.device)
    if isinstance(gamma_mn, (int, np.integer)):
        baryon_two_point(out, propag_i, propag_j, propag_n, contract_type, V, gamma_ij, gamma_kl, gamma_mn)
    else:
        baryon_general_two_point(out, propag_i, propag_j, propag_n, contract_type, V, gamma_ij, gamma_kl,
                                 np.asarray(gamma_mn))
    return out


def baryonSequentialTwoPoint(propag_i, propag_j, propag_n, contract_type, sequential_type, gamma_ij: int,
                             gamma_kl: int, gamma_mn: int):
    """Overwrites and returns the propagator selected by ``sequential_type``."""
    V = _fp64_volume(propag_i, propag_j, propag_n)
    baryon_sequential_two_point(propag_i, propag_j, propag_n, contract_type, sequential_type, V, gamma_ij,
                                gamma_kl, gamma_mn)
    return (propag_i, propag_j, propag_n)[int(sequential_type)]


_GAMMA_CACHE: dict = {}
_PROJ_CACHE: dict = {}


def _projectors(insertions) -> np.ndarray:
    """[n][4][4] complex128 projectors; gamma-index lists are cached so the
    per-call host cost stays small next to the kernel."""
    key = tuple(insertions) if all(isinstance(w, (int, np.integer)) for w in insertions) else None
    if key is not None and key in _PROJ_CACHE:
        return _PROJ_CACHE[key]
    mats = []
    for w in insertions:
        if isinstance(w, (int, np.integer)):
            g = int(w)
            if g not in _GAMMA_CACHE:
                _GAMMA_CACHE[g] = gamma_matrix(g)
            mats.append(_GAMMA_CACHE[g])
        else:
            mats.append(np.asarray(w, dtype=np.complex128).reshape(4, 4))
    out = np.ascontiguousarray(np.stack(mats))
    if key is not None:
        out.setflags(write=False)
        _PROJ_CACHE[key] = out
    return out


def project_workspace_bytes(dims: Sequence[int], nmom: int, site_order: int = LEXICOGRAPHIC) -> int:
    d = np.ascontiguousarray(dims, dtype=np.int32)
    return int(lib().contract_project_workspace_bytes(d.ctypes.data_as(ctypes.c_void_p), int(site_order),
                                                      int(nmom)))


class Sublattice(ctypes.Structure):
    """contract_sublattice."""
    _fields_ = [("dims", ctypes.c_int * 4), ("global_", ctypes.c_int * 4), ("origin", ctypes.c_int * 4),
                ("site_order", ctypes.c_int)]


def baryonTwoPointProject(propag_i, propag_j, propag_n, terms, gamma_ij: int, gamma_kl: int, insertions,
                          momenta, dims, site_order: int = LEXICOGRAPHIC, out=None, workspace=None,
                          stream=None, global_dims=None, origin=None):
    """Fused baryon two-point: sum over ``terms`` = [(contract_type, weight)],
    for every insertion (gamma index or 4x4 projector), projected to
    ``out[len(insertions), len(momenta), Lt]`` (complex128, on the device).

    Propagators may be complex128 (FP64 path), complex64 (FP32 contraction,
    FP64 accumulation) or ScaledPropagator (FP32 site-scaled form, see
    toScaledFP32).  global_dims / origin place the local lattice inside a
    larger one (one rank of an MPI grid): phases use global coordinates and
    momenta are in units of 2pi / global extent."""
    import torch
    precision = _precision((propag_i, propag_j, propag_n))
    _check_lattice_props((propag_i, propag_j, propag_n), dims, "baryonTwoPointProject")
    W = _projectors(insertions)
    mom = np.ascontiguousarray(momenta, dtype=np.int32).reshape(-1, 3)
    types = np.ascontiguousarray([int(t) for t, _ in terms], dtype=np.int32)
    weights = np.ascontiguousarray([complex(w) for _, w in terms], dtype=np.complex128)
    d = np.ascontiguousarray(dims, dtype=np.int32)
    if out is None:
        out = torch.empty((W.shape[0], mom.shape[0], int(d[3])), dtype=torch.complex128, device=propag_i.device)
    _check_out(out, (W.shape[0], mom.shape[0], int(d[3])))
    _check_workspace(workspace)
    _same_device((propag_i, propag_j, propag_n, out, workspace), "propagators, output and workspace")
    ws_ptr, ws_bytes = 0, 0
    if workspace is not None:
        ws_ptr, ws_bytes = workspace.data_ptr(), workspace.numel() * workspace.element_size()
    vp = ctypes.c_void_p
    if global_dims is None and origin is None:
        rc = lib().contract_baryon_twopoint_project(
            _ptr(out), _ptr(propag_i), _ptr(propag_j), _ptr(propag_n), precision, types.ctypes.data_as(vp),
            weights.ctypes.data_as(vp), len(types), int(gamma_ij), int(gamma_kl), W.ctypes.data_as(vp), W.shape[0],
            mom.ctypes.data_as(vp), mom.shape[0], d.ctypes.data_as(vp), int(site_order), ws_ptr, ws_bytes,
            0 if stream is None else int(stream))
    else:
        lat = Sublattice()
        for mu in range(4):
            lat.dims[mu] = int(d[mu])
            lat.global_[mu] = int(global_dims[mu]) if global_dims is not None and mu < len(global_dims) else 0
            lat.origin[mu] = int(origin[mu]) if origin is not None and mu < len(origin) else 0
        lat.site_order = int(site_order)
        rc = lib().contract_baryon_twopoint_project_sub(
            _ptr(out), _ptr(propag_i), _ptr(propag_j), _ptr(propag_n), precision, types.ctypes.data_as(vp),
            weights.ctypes.data_as(vp), len(types), int(gamma_ij), int(gamma_kl), W.ctypes.data_as(vp), W.shape[0],
            mom.ctypes.data_as(vp), mom.shape[0], ctypes.byref(lat), ws_ptr, ws_bytes,
            0 if stream is None else int(stream))
    if rc != 0:
        raise RuntimeError(f"contract_baryon_twopoint_project failed ({rc}): {last_error()}")
    return out


class BaryonRequest(ctypes.Structure):
    """contract_baryon_request."""
    _fields_ = [("propag_i", ctypes.c_void_p), ("propag_j", ctypes.c_void_p), ("propag_n", ctypes.c_void_p),
                ("out", ctypes.c_void_p), ("lattice", Sublattice)]


def _baryon_requests(requests):
    """Request tuples (Pi, Pj, Pn, dims[, site_order[, global_dims[, origin]]])
    as a contract_baryon_request array."""
    arr = (BaryonRequest * len(requests))()
    for r, req in zip(arr, requests):
        r.propag_i, r.propag_j, r.propag_n = _ptr(req[0]), _ptr(req[1]), _ptr(req[2])
        for mu in range(4):
            r.lattice.dims[mu] = int(req[3][mu])
        r.lattice.site_order = int(req[4]) if len(req) > 4 else LEXICOGRAPHIC
        gl = req[5] if len(req) > 5 and req[5] is not None else ()
        og = req[6] if len(req) > 6 and req[6] is not None else ()
        for mu in range(4):
            r.lattice.global_[mu] = int(gl[mu]) if mu < len(gl) else 0
            r.lattice.origin[mu] = int(og[mu]) if mu < len(og) else 0
    return arr


def baryon_batch_workspace_bytes(requests, nmom: int) -> int:
    arr = _baryon_requests(requests)
    return int(lib().contract_baryon_batch_workspace_bytes(ctypes.cast(arr, ctypes.c_void_p), len(arr), int(nmom)))


def baryonTwoPointProjectBatch(requests, terms, gamma_ij: int, gamma_kl: int, insertions, momenta, outs=None,
                               workspace=None, stream=None):
    """Several fused baryon two-points in one launch.  ``requests`` holds
    tuples (Pi, Pj, Pn, dims[, site_order[, global_dims[, origin]]]); terms,
    gammas, insertions and momenta are shared.  Returns one
    ``[len(insertions), len(momenta), Lt]`` complex128 tensor per request,
    each equal to what baryonTwoPointProject gives for that request alone."""
    import torch
    if not requests:
        return []
    precision = _precision([p for req in requests for p in req[:3]])
    for k, req in enumerate(requests):
        _check_lattice_props(req[:3], req[3], f"request {k}")
    W = _projectors(insertions)
    mom = np.ascontiguousarray(momenta, dtype=np.int32).reshape(-1, 3)
    types = np.ascontiguousarray([int(t) for t, _ in terms], dtype=np.int32)
    weights = np.ascontiguousarray([complex(w) for _, w in terms], dtype=np.complex128)
    arr = _baryon_requests(requests)
    if outs is None:
        outs = [torch.empty((W.shape[0], mom.shape[0], int(req[3][3])), dtype=torch.complex128,
                            device=req[0].device) for req in requests]
    if len(outs) != len(requests):
        raise ValueError(f"{len(outs)} outputs for {len(requests)} requests")
    for k, (req, o) in enumerate(zip(requests, outs)):
        _check_out(o, (W.shape[0], mom.shape[0], int(req[3][3])), f"outs[{k}]")
    _check_workspace(workspace)
    _same_device([p for req in requests for p in req[:3]] + list(outs) + [workspace],
                 "propagators, outputs and workspace")
    for r, o in zip(arr, outs):
        r.out = _ptr(o)
    ws_ptr, ws_bytes = 0, 0
    if workspace is not None:
        ws_ptr, ws_bytes = workspace.data_ptr(), workspace.numel() * workspace.element_size()
    vp = ctypes.c_void_p
    rc = lib().contract_baryon_twopoint_project_batch(
        ctypes.cast(arr, vp), len(arr), precision, types.ctypes.data_as(vp), weights.ctypes.data_as(vp),
        len(types), int(gamma_ij), int(gamma_kl), W.ctypes.data_as(vp), W.shape[0], mom.ctypes.data_as(vp),
        mom.shape[0], ws_ptr, ws_bytes, 0 if stream is None else int(stream))
    if rc != 0:
        raise RuntimeError(f"contract_baryon_twopoint_project_batch failed ({rc}): {last_error()}")
    return outs


# ---------------------------------------------------------------------------
# Three-point step after the sequential source (include/contract_ext.h)
# ---------------------------------------------------------------------------

class ThreePointRequest(ctypes.Structure):
    """contract_threepoint_request."""
    _fields_ = [("propag_fwd", ctypes.c_void_p), ("propag_seq", ctypes.c_void_p), ("out", ctypes.c_void_p),
                ("lattice", Sublattice)]


def _threepoint_requests(requests):
    """Request tuples (F, Sigma, dims[, site_order[, global_dims[, origin]]])."""
    arr = (ThreePointRequest * len(requests))()
    for r, req in zip(arr, requests):
        r.propag_fwd, r.propag_seq = _ptr(req[0]), _ptr(req[1])
        for mu in range(4):
            r.lattice.dims[mu] = int(req[2][mu])
        r.lattice.site_order = int(req[3]) if len(req) > 3 else LEXICOGRAPHIC
        gl = req[4] if len(req) > 4 and req[4] is not None else ()
        og = req[5] if len(req) > 5 and req[5] is not None else ()
        for mu in range(4):
            r.lattice.global_[mu] = int(gl[mu]) if mu < len(gl) else 0
            r.lattice.origin[mu] = int(og[mu]) if mu < len(og) else 0
    return arr


def threepoint_workspace_bytes(requests, nmom: int) -> int:
    arr = _threepoint_requests(requests)
    return int(lib().contract_threepoint_workspace_bytes(ctypes.cast(arr, ctypes.c_void_p), len(arr), int(nmom)))


def sequentialSource(seq):
    """The solver source sigma of a sequential source seq = dC/dP_X (what
    baryon_sequential_two_point leaves in P_X's buffer):
    sigma[s1][s2](c1,c2) = g5[s1] conj(seq[s1][s2](c1,c2)).  Solving
    D Sigma = sigma gives the sequential propagator of threePointProject.
    A ScaledPropagator gives a ScaledPropagator (the transformation only
    conjugates and flips signs, so the site exponents carry over)."""
    import torch
    if isinstance(seq, ScaledPropagator):
        out = seq.clone()
        m = out.entries()
        g5 = torch.tensor([1.0, 1.0, -1.0, -1.0], dtype=torch.float32, device=m.device)
        m.copy_(m.conj() * g5.view(1, 4, 1, 1, 1))
        return out
    g5 = torch.tensor([1.0, 1.0, -1.0, -1.0], dtype=seq.real.dtype, device=seq.device)
    return (seq.conj() * g5.view(1, 4, 1, 1, 1)).contiguous()


def baryonSequential(propag_i, propag_j, propag_n, contract_type, sequential_type, gamma_ij: int, gamma_kl: int,
                     gamma_mn: int, stream=None):
    """contract_baryon_sequential: the sequential source seq_X = dC/dP_X into
    the propagator selected by ``sequential_type`` (overwritten, never read),
    which is returned.  All three complex128 tensors (FP64, as
    baryon_sequential_two_point) or all three ScaledPropagator (FP32
    site-scaled storage in and out)."""
    props = (propag_i, propag_j, propag_n)
    scaled = [isinstance(p, ScaledPropagator) for p in props]
    if any(scaled) and not all(scaled):
        raise TypeError("the three propagators must share one storage form")
    if all(scaled):
        V = props[0].volume
        for p in props:
            if _volume(p) != V:
                raise ValueError(f"propagators hold different site counts {[q.volume for q in props]}")
        _same_device(props, "propagators")
        precision = CONTRACT_FP32_SCALED
    else:
        V = _fp64_volume(*props)
        precision = CONTRACT_FP64
    rc = lib().contract_baryon_sequential(_ptr(propag_i), _ptr(propag_j), _ptr(propag_n), precision,
                                          int(contract_type), int(sequential_type), V, int(gamma_ij), int(gamma_kl),
                                          int(gamma_mn), 0 if stream is None else int(stream))
    if rc != 0:
        raise RuntimeError(f"contract_baryon_sequential failed ({rc}): {last_error()}")
    return props[int(sequential_type)]


def baryonDiquarkInto(out, propag_i, propag_j, gamma_ij: int, gamma_kl: int, stream=None):
    """contract_baryon_diquark: the diquark of propag_i and propag_j into
    ``out``, which is returned (baryonDiquark allocates a complex128 output
    instead).  Three complex128 tensors (FP64, as
    baryon_diquark) or three ScaledPropagator (FP32 site-scaled storage in
    and out)."""
    props = (out, propag_i, propag_j)
    scaled = [isinstance(p, ScaledPropagator) for p in props]
    if any(scaled) and not all(scaled):
        raise TypeError("the output and both propagators must share one storage form")
    if all(scaled):
        V = props[0].volume
        for p in props:
            if _volume(p) != V:
                raise ValueError(f"propagators hold different site counts {[q.volume for q in props]}")
        _same_device(props, "propagators")
        precision = CONTRACT_FP32_SCALED
    else:
        V = _fp64_volume(*props)
        precision = CONTRACT_FP64
    rc = lib().contract_baryon_diquark(_ptr(out), _ptr(propag_i), _ptr(propag_j), precision, V, int(gamma_ij),
                                       int(gamma_kl), 0 if stream is None else int(stream))
    if rc != 0:
        raise RuntimeError(f"contract_baryon_diquark failed ({rc}): {last_error()}")
    return out


def threePointProject(requests, insertions, momenta, outs=None, workspace=None, stream=None):
    """Three-point functions after the sequential source: for each request
    (F, Sigma, dims[, site_order[, global_dims[, origin]]]) return
    out[len(insertions)][len(momenta)][Lt] (complex128, on the device) with
    out[w][p][t] = sum_{x in t} exp(-i p.x) Tr[g5 Gamma_w F(x) Sigma(x)^dagger],
    Gamma_w a gamma index or a 4x4 matrix.  F and Sigma are complex128,
    complex64 or ScaledPropagator (one storage form per call)."""
    import torch
    if not requests:
        return []
    precision = _precision([p for req in requests for p in req[:2]])
    for k, req in enumerate(requests):
        _check_lattice_props(req[:2], req[2], f"request {k}")
    G = _projectors(insertions)
    mom = np.ascontiguousarray(momenta, dtype=np.int32).reshape(-1, 3)
    if outs is None:
        outs = [torch.empty((G.shape[0], mom.shape[0], int(req[2][3])), dtype=torch.complex128, device=req[0].device)
                for req in requests]
    if len(outs) != len(requests):
        raise ValueError(f"{len(outs)} outputs for {len(requests)} requests")
    for k, (req, o) in enumerate(zip(requests, outs)):
        _check_out(o, (G.shape[0], mom.shape[0], int(req[2][3])), f"outs[{k}]")
    _check_workspace(workspace)
    _same_device([p for req in requests for p in req[:2]] + list(outs) + [workspace],
                 "propagators, outputs and workspace")
    arr = _threepoint_requests(requests)
    for r, o in zip(arr, outs):
        r.out = _ptr(o)
    ws_ptr, ws_bytes = 0, 0
    if workspace is not None:
        ws_ptr, ws_bytes = workspace.data_ptr(), workspace.numel() * workspace.element_size()
    vp = ctypes.c_void_p
    rc = lib().contract_threepoint_project(ctypes.cast(arr, vp), len(arr), precision, G.ctypes.data_as(vp),
                                           G.shape[0], mom.ctypes.data_as(vp), mom.shape[0], ws_ptr, ws_bytes,
                                           0 if stream is None else int(stream))
    if rc != 0:
        raise RuntimeError(f"contract_threepoint_project failed ({rc}): {last_error()}")
    return outs


# ---------------------------------------------------------------------------
# Batched meson Gamma x Gamma' grids (include/contract_ext.h)
# ---------------------------------------------------------------------------

class MesonRequest(ctypes.Structure):
    """contract_meson_request."""
    _fields_ = [("propag_i", ctypes.c_void_p), ("propag_j", ctypes.c_void_p), ("out", ctypes.c_void_p),
                ("dims", ctypes.c_int * 4), ("site_order", ctypes.c_int), ("global_", ctypes.c_int * 3),
                ("origin", ctypes.c_int * 3)]


def _fill_geometry(r, req):
    """Request tuple (Pi, Pj, dims[, site_order[, global_dims[, origin]]])."""
    for mu in range(4):
        r.dims[mu] = int(req[2][mu])
    r.site_order = int(req[3]) if len(req) > 3 else LEXICOGRAPHIC
    gl = req[4] if len(req) > 4 and req[4] is not None else (0, 0, 0)
    og = req[5] if len(req) > 5 and req[5] is not None else (0, 0, 0)
    for mu in range(3):
        r.global_[mu] = int(gl[mu])
        r.origin[mu] = int(og[mu])


class MesonGridTuning(ctypes.Structure):
    """contract_meson_grid_tuning; zero fields take the library default.
    trace_unit: 1 vector ALUs, 2 matrix cores (FP64 / site-scaled FP32), 3 matrix cores
    with one wave per 4-site tile holding all of K (passes of at most 4 momentum slots)."""
    _fields_ = [("sites_per_item", ctypes.c_int), ("tile_sites", ctypes.c_int), ("blocks_per_cu", ctypes.c_int),
                ("trace_unit", ctypes.c_int)]

    def key(self):
        return (self.sites_per_item, self.tile_sites, self.blocks_per_cu, self.trace_unit)


def _tuning(tuning):
    if tuning is None:
        return None
    if isinstance(tuning, MesonGridTuning):
        return tuning
    return MesonGridTuning(*[int(v) for v in tuning])


def _request_array(requests, outs):
    arr = (MesonRequest * len(requests))()
    for r, (req, out) in enumerate(zip(requests, outs)):
        arr[r].propag_i = _ptr(req[0])
        arr[r].propag_j = _ptr(req[1])
        arr[r].out = _ptr(out)
        _fill_geometry(arr[r], req)
    return arr


def meson_grid_workspace_bytes(requests, nmom: int, tuning=None) -> int:
    """Device scratch (bytes) of mesonGridProject for these requests; a request
    is (propag_i, propag_j, dims[, site_order]), only dims/site_order are read."""
    arr = (MesonRequest * len(requests))()
    for r, req in enumerate(requests):
        arr[r].propag_i = arr[r].propag_j = arr[r].out = 1
        _fill_geometry(arr[r], req)
    t = _tuning(tuning)
    return int(lib().contract_meson_grid_workspace_bytes(arr, len(requests), int(nmom),
                                                         ctypes.byref(t) if t is not None else None))


def mesonGridProject(requests, momenta, tuning=None, outs=None, workspace=None, stream=None):
    """Batched meson two-point grids: for each request (propag_i, propag_j,
    dims[, site_order[, global_dims[, origin]]]) return out[16 Gamma_ij][16 Gamma_kl][len(momenta)][Lt]
    (complex128 on the device) with out[g1][g2][p][t] = sum_{x in t}
    exp(-i p.x) meson_two_point(propag_i, propag_j; g1, g2)(x).

    All requests go through one persistent launch (per 32 requests and 16
    +-k momentum slots, a momentum and its negative sharing a slot).  complex64 propagators (or ScaledPropagator, see toScaledFP32)
    select the FP32-contraction / FP64-accumulation form; all requests must
    share the storage form."""
    import torch
    if not requests:
        return []
    precision = _precision([p for req in requests for p in req[:2]])
    for k, req in enumerate(requests):
        _check_lattice_props(req[:2], req[2], f"request {k}")
    mom = np.ascontiguousarray(momenta, dtype=np.int32).reshape(-1, 3)
    if outs is None:
        outs = [torch.empty((16, 16, mom.shape[0], int(req[2][3])), dtype=torch.complex128, device=req[0].device)
                for req in requests]
    if len(outs) != len(requests):
        raise ValueError(f"{len(outs)} outputs for {len(requests)} requests")
    for k, (req, o) in enumerate(zip(requests, outs)):
        _check_out(o, (16, 16, mom.shape[0], int(req[2][3])), f"outs[{k}]")
    _check_workspace(workspace)
    _same_device([p for req in requests for p in req[:2]] + list(outs) + [workspace],
                 "propagators, outputs and workspace")
    arr = _request_array(requests, outs)
    t = _tuning(tuning)
    ws_ptr, ws_bytes = 0, 0
    if workspace is not None:
        ws_ptr, ws_bytes = workspace.data_ptr(), workspace.numel() * workspace.element_size()
    rc = lib().contract_meson_grid_project(arr, len(requests), precision, mom.ctypes.data_as(ctypes.c_void_p),
                                           mom.shape[0], ctypes.byref(t) if t is not None else None, ws_ptr,
                                           ws_bytes, 0 if stream is None else int(stream))
    if rc != 0:
        raise RuntimeError(f"contract_meson_grid_project failed ({rc}): {last_error()}")
    return outs


class BaryonPlan(ctypes.Structure):
    """contract_baryon_plan; zero fields take the library default."""
    _fields_ = [("items_per_cu", ctypes.c_int), ("min_item_sites", ctypes.c_int)]

    def key(self):
        return (self.items_per_cu, self.min_item_sites)


def get_baryon_plan():
    p = BaryonPlan()
    if lib().contract_get_baryon_plan(ctypes.byref(p)) != 0:
        raise RuntimeError(last_error())
    return p.key()


def set_baryon_plan(plan) -> None:
    """(items_per_cu, min_item_sites) for this thread's fused baryon calls;
    None restores the defaults."""
    if plan is None:
        rc = lib().contract_set_baryon_plan(None)
    else:
        p = BaryonPlan(*[int(v) for v in plan])
        rc = lib().contract_set_baryon_plan(ctypes.byref(p))
    if rc != 0:
        raise ValueError(f"contract_set_baryon_plan failed: {last_error()}")


class baryon_plan:
    """Context manager: the fused baryon entry points called from this
    thread use work plan ``plan`` = (items_per_cu, min_item_sites)."""

    def __init__(self, plan):
        self.plan = plan

    def __enter__(self):
        self.saved = get_baryon_plan()
        set_baryon_plan(self.plan)
        return self

    def __exit__(self, *exc):
        set_baryon_plan(self.saved)
        return False


class abi_stream:
    """Context manager: drop-in entry points called from this thread launch on
    ``stream`` (a torch stream, a raw handle or None for the null stream)."""

    def __init__(self, stream):
        self.handle = getattr(stream, "cuda_stream", stream) or 0

    def __enter__(self):
        lib().contract_set_stream(self.handle)
        return self

    def __exit__(self, *exc):
        lib().contract_set_stream(None)
        return False
