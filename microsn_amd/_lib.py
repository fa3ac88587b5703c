"""ctypes binding of libmlc.so (the C ABI declared in include/mlc.h and
include/mlc_host.h).

The library is built in-tree (``make -C microsn_amd/csrc``).  Loading never
falls back to anything else: if ``libmlc.so`` is missing the import of this
module's ``lib()`` raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmlc.so")

F32, F64 = 0, 1
SHAPES = {"uniform": 0, "gaussian": 1, "gaussian_hole": 2, "thermal_hole": 3, "wavy": 4,
          "exponential": 5}

# Every exported symbol with (restype, argtypes).  Kept in step with the
# headers by tests/test_abi.py.
_i, _d, _p, _i64, _sz = C.c_int, C.c_double, C.c_void_p, C.c_int64, C.c_size_t
_pp = C.POINTER(C.c_void_p)
SIGNATURES = {
    # mlc.h: context
    "mlc_version": (_i, []),
    "mlc_ctx_create": (_i, [_i, _p, _pp]),
    "mlc_ctx_destroy": (None, [_p]),
    "mlc_last_error": (C.c_char_p, [_p]),
    "mlc_ctx_set_stream": (_i, [_p, _p]),
    "mlc_ctx_generation": (_i, [_p, C.POINTER(C.c_uint64)]),
    "mlc_ctx_reserve": (_i, [_p, _sz]),
    "mlc_ctx_sync": (_i, [_p]),
    "mlc_malloc": (_i, [_p, _sz, _pp]),
    "mlc_free": (_i, [_p, _p]),
    "mlc_memcpy_h2d": (_i, [_p, _p, _p, _sz]),
    "mlc_memcpy_d2h": (_i, [_p, _p, _p, _sz]),
    "mlc_memset": (_i, [_p, _p, _i, _sz]),
    "mlc_mem_info": (_i, [_p, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "mlc_timing_enable": (_i, [_p, _i]),
    "mlc_timing_reset": (_i, [_p]),
    "mlc_timing_get": (_i, [_p, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
    "mlc_timing_names": (_i, [_p, C.c_char_p, _sz]),
    # maps and convolution
    "mlc_gerlumph_normalise": (_i, [_p, _p, _i64, _d, _d, _d, _d, _p]),
    "mlc_map_create": (_i, [_p, _i, _p, _i, _i, _i, _pp]),
    "mlc_map_destroy": (None, [_p]),
    "mlc_map_update": (_i, [_p, _p]),
    "mlc_profile": (_i, [_p, _i, _d, _d, _d, _d, C.POINTER(C.c_int), _p, _p, _i64]),
    "mlc_parity_round": (_d, [_d]),
    "mlc_profile_size": (_i, [_i, _d, _d, _d, _d]),
    "mlc_map_convolve": (_i, [_p, _i, _d, _d, _d, _d, _p, _p]),
    "mlc_map_convolve_nodes": (_i, [_p, _i, _p, _p, _d, _p, _p, _p]),
    "mlc_maps_update": (_i, [_i, _p, _p]),
    "mlc_maps_convolve_nodes": (_i, [_i, _p, _p, _p, _p, _d, _p, _p, _p]),
    "mlc_map_convolve_plain": (_i, [_p, _i, _d, _d, _d, _d, _p, _p]),
    "mlc_fft_r2c": (_i, [_p, _i, _p, _p, _i, _i]),
    "mlc_fft_c2r": (_i, [_p, _i, _p, _p, _i, _i]),
    "mlc_sample_points": (_i, [_p, _i, _p, _i, _i, _i, _i, _i, _p, _p, _p, _p]),
    "mlc_expanding_source": (_i, [_p, _i, _p, _i, _i, _d, _p, _i, _i, _i, _p, _p, _p, _p]),
    # sort / pair mode / MPD
    "mlc_sort_f64": (_i, [_p, _p, _i64, _p, _p, _p]),
    "mlc_bin_edges": (_i, [_p, _p, _i64, _p, _i, _p, _p]),
    "mlc_setup_integral": (_i, [_p, _p, _p, _i, _i, _p, _i, _p, _p, _p]),
    "mlc_calculate_chi2": (_i, [_p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _p]),
    "mlc_bin_chi2": (_i, [_p, _p, _i64, _p, _p, _i, _p, _p]),
    "mlc_bin_table_device": (_i, [_p, _p, _i64, _p, _i, _p, _p, _p, _p]),
    "mlc_binned_mpd": (_i, [_p, _p, _i64, _i, _d, _d, _p, _p]),
    "mlc_chi2_multiband": (_i, [_p, _p, _p, _p, _p, _i, _i, _i, _i, _p, _p]),
    # light curves
    "mlc_lc_chi2_grad": (_i, [_p, _i, _p, _i, _i, _i, _i, _i, _p, _p, _p, _p, _i, _p, _p, _p,
                              _i, _p, _i, _p, _p, _p, _p, _p]),
    "mlc_lc_chi2_grad_convolve": (_i, [_p, _p, _d, _p, _i, _i, _i, _p, _p, _p, _p, _i, _p, _p, _p, _i,
                                       _p, _i, _p, _p, _p, _p, _p]),
    "mlc_lc_chi2_grad_batch": (_i, [_p, _i, _i, _p, _p, _i, _p, _p, _i, _i, _p, _p, _p, _i, _p, _p, _p, _p,
                                    _i, _p, _i, _p, _p, _p, _p, _p]),
    "mlc_tune_mode": (_i, [_p, _i]),
    "mlc_tune_force": (_i, [_p, _i, _i, _i]),
    "mlc_tune_lookup": (_i, [_p, _i, _i, _d, _d, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "mlc_tune_save": (_i, [_p, C.c_char_p]),
    "mlc_tune_load": (_i, [_p, C.c_char_p]),
    "mlc_lc_chi2_grad_noalloc": (_i, [_p, _i, _p, _i, _i, _i, _i, _i, _p, _p, _p, _p, _i, _p, _p,
                                      _p, _i, _p, _i, _p, _p, _p, _p, _p]),
    "mlc_lc_chi2_grad_plain": (_i, [_p, _i, _p, _i, _i, _i, _i, _i, _p, _p, _p, _p, _i, _p, _p,
                                    _p, _i, _p, _i, _p, _p, _p, _p, _p]),
    # mlc_host.h
    "mlc_locations_on_grid": (_i, [_i, _i, _i, _p, _p]),
    "mlc_locations_random": (_i, [_i, _i, _i, _i, _p, _p]),
    "mlc_find_indices": (_i, [_p, _p, _i, _i, _p]),
    "mlc_setup_chi2": (_i, [_d, _d, _d, _d, _i, _p, _p, _p, _i, _p, _i, _p, _p, _p, _p, _p, _p,
                            C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "mlc_mpd_bins": (_i, [_i, _d, _d, _i, _p]),
    "mlc_mpd_normalize": (_i, [_i, _p, _p]),
    "mlc_mpd_divide_by": (_i, [_i, _p, _p, _i, _p, _p, _i, _p, _p]),
    "mlc_trapezium": (_d, [_i, _p, _p, _p]),
    "mlc_gerlumph_read": (_i, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                               C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double),
                               C.POINTER(C.c_double), C.POINTER(C.c_double), _p, _i64]),
    "mlc_gerlumph_write": (_i, [C.c_char_p, _d, _d, _i, _d, _d, _d, _d, _p]),
    "mlc_write_fits": (_i, [C.c_char_p, _p, _i, _i, _i]),
    "mlc_read_fits": (_i, [C.c_char_p, _i, _p, _i64, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "mlc_write_bin": (_i, [C.c_char_p, _p, _i, _i, _i]),
    "mlc_size_parametric": (_d, [_d, _d, _d, _d]),
    "mlc_size_ss": (_d, [_d, _d, _d, _d]),
    "mlc_size_parametric_band": (_d, [_d, _d, _d, _p, _p, _i]),
    "mlc_size_ss_band": (_d, [_d, _d, _d, _p, _p, _i]),
    "mlc_size_vector_band": (_d, [_p, _p, _i]),
}

_lib = None


class MlcError(RuntimeError):
    """A non-zero status from libmlc.so, with the context's last error."""

    def __init__(self, code, msg):
        super().__init__(f"libmlc status {code}: {msg}")
        self.code = code


def lib():
    """The loaded library.  Raises if libmlc.so has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C microsn_amd/csrc` "
                              "(or __graft_entry__.build())")
        # torch ships its own HIP runtime; importing it first makes libmlc.so
        # bind to that one instead of loading a second runtime into the process
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(x):
    """Raw address of a torch tensor, numpy array, int or None."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    if hasattr(x, "ctypes"):
        return x.ctypes.data
    raise TypeError(f"cannot take the address of {type(x)}")
