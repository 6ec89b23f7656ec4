"""coqui_amd: MI355X-native THC imaginary-time GW/RPA inner loop.

The compute lives in libthcgw.so (hand-written HIP kernels for gfx950 behind the C ABI of
include/thcgw.h).  This package holds the Python binding (thcgw.ThcGW), the IR tables
(iaft.IAFT), the seeded synthetic-system generator (system.make_system) and the
self-consistent loop (scgw.ScGW) with the learned Krylov Dyson solve (thcgw.LearnedDyson).
"""
from .iaft import IAFT, available_tables, determine_lambda  # noqa: F401
from .system import System, make_system  # noqa: F401
from .thcgw import EXPORTS, LIB_PATH, LearnedDyson, Mlp, ThcBatch, ThcGW, ThcgwError, load_library  # noqa: F401
from .scgw import ScGW  # noqa: F401
from .eri import read_thc_eri, write_thc_eri  # noqa: F401

__all__ = ["IAFT", "available_tables", "determine_lambda", "System", "make_system", "ThcGW", "ThcBatch", "ThcgwError",
           "LearnedDyson", "Mlp", "ScGW", "read_thc_eri", "write_thc_eri",
           "load_library", "EXPORTS", "LIB_PATH"]
