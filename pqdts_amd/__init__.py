"""pqdts_amd: MI355X-native two-metric projected truncated-Newton solver for
quantum detector tomography (a drop-in for pc2/pqdts's pqdts_omp.x).

The compute path is libpqdts_amd.so (HIP kernels for gfx950 behind the C ABI
in include/pqdts_amd.h). Python is a thin ctypes layer over it.
"""
from ._lib import (COL_MAJOR, ROW_MAJOR, SLOTS, Batch, Context, PqdtsError, Server, device_count,  # noqa: F401
                   device_cu_count, footprint, load, simplexproj_condat, simplexproj_rows)

__all__ = ["Batch", "Context", "Server", "footprint", "PqdtsError", "device_count", "device_cu_count", "load", "simplexproj_rows", "simplexproj_condat", "SLOTS", "ROW_MAJOR", "COL_MAJOR"]
