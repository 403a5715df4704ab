"""spec_spars_amd -- MI355X (gfx950) sparsification hot path of spec_spars.

The compute path is the HIP library libspecspars_hip.so behind the C ABI of
include/spec_spars_hip.h; this package is its Python binding (ctypes) plus
synthetic graph generators.  Importing the package loads the library and
fails if it has not been built.
"""
from . import _lib
from ._lib import (SSError, PREC_FP64, PREC_MIXED, STOP_REF_ABS, STOP_REL,  # noqa: F401
                   ORDER_NATURAL, ORDER_RANDOM, ORDER_ROUNDS, MIXED_RELIABLE, MIXED_RESTART,
                   MIXED_RELIABLE_F16)

_lib.load()

from .api import (Batch, Context, Factor, Graph, set_factor_threads, solve_opts, mixed_mode_for,  # noqa: E402,F401
                  FP16_ROWS_MAX_SPREAD)
from . import graphs  # noqa: E402,F401
from .sparsifier import InputStream, SparsifierParameters, Sparsifier, read_mtx, write_mtx  # noqa: E402,F401
