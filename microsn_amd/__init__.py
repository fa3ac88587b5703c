"""microsn_amd -- the microLC light-curve simulation + chi2 path on MI355X.

All computation runs in the in-tree HIP library ``libmlc.so`` (sources in
``csrc/``, C ABI in ``include/mlc.h`` and ``include/mlc_host.h``); this
package binds it with ctypes over torch device memory.
"""
from ._lib import F32, F64, SHAPES, MlcError, lib  # noqa: F401

__all__ = ["F32", "F64", "SHAPES", "MlcError", "lib"]
