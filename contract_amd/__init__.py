"""contract_amd -- MI355X (gfx950) propagator-contraction library.

The product is ``contract_amd/lib/libcontract.so``: the drop-in C ABI of
include/contract.h plus the fused timeslice/momentum extension of
include/contract_ext.h, with hand-written HIP kernels in ``csrc/``.
``contract_amd.pycontract`` is its Python face.
"""

from ._lib import ABI_SYMBOLS, EXT_SYMBOLS, LIB_PATH, LibraryMissing, last_error, lib  # noqa: F401
from . import pycontract  # noqa: F401
