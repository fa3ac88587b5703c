"""The device-independent cores of the gfx950 kernels, compiled for the CPU
with g++ from the same headers: the Stockham butterflies and real-FFT
pre/post steps of mlc_fft_core.hpp against numpy's FFT, and the FP64
logarithm of mlc_math.hpp against numpy's log.  CPU only."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _build(src, tmp_path_factory):
    out = tmp_path_factory.mktemp("harness") / (os.path.splitext(src)[0] + ".so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                           "-o", str(out), os.path.join(HERE, src)])
    return C.CDLL(str(out))


@pytest.fixture(scope="module")
def fftlib(tmp_path_factory):
    L = _build("fft_core_harness.cpp", tmp_path_factory)
    L.h_fft.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    L.h_rfft.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.h_irfft.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.h_ctfft.argtypes = [C.c_void_p, C.c_int, C.c_int]
    return L


@pytest.fixture(scope="module")
def mathlib(tmp_path_factory):
    L = _build("math_harness.cpp", tmp_path_factory)
    L.h_ln.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.h_ln_f32.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


LENGTHS = [2, 3, 4, 5, 8, 15, 16, 32, 60, 64, 100, 128, 240, 256, 1000, 1024, 2048, 2500, 4096, 5000]


@pytest.mark.parametrize("L", LENGTHS)
@pytest.mark.parametrize("inverse", [0, 1])
@pytest.mark.parametrize("quarter", [0, 1])
def test_complex_fft_matches_numpy(fftlib, L, inverse, quarter):
    rng = np.random.default_rng(L)
    x = rng.standard_normal(L) + 1j * rng.standard_normal(L)
    buf = np.ascontiguousarray(np.stack([x.real, x.imag], 1))
    assert fftlib.h_fft(_p(buf), L, inverse, quarter) == 0
    got = buf[:, 0] + 1j * buf[:, 1]
    want = L * np.fft.ifft(x) if inverse else np.fft.fft(x)
    assert np.max(np.abs(got - want)) < 1e-13 * np.sqrt(L) * np.max(np.abs(want))


@pytest.mark.parametrize("N", [256, 512, 1024, 2048, 4096, 8192, 16384])
@pytest.mark.parametrize("inverse", [0, 1])
def test_compile_time_fft_matches_numpy(fftlib, N, inverse):
    """mlc_ctfft.hpp (radix-16 Stockham, two-level twiddle tables, padded
    LDS indices), run thread by thread around each barrier."""
    rng = np.random.default_rng(N + inverse)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    buf = np.ascontiguousarray(np.stack([x.real, x.imag], 1))
    assert fftlib.h_ctfft(_p(buf), N, inverse) == 0
    got = buf[:, 0] + 1j * buf[:, 1]
    want = N * np.fft.ifft(x) if inverse else np.fft.fft(x)
    assert np.max(np.abs(got - want)) < 1e-13 * np.sqrt(N) * np.max(np.abs(want))


def test_unsupported_length_is_rejected(fftlib):
    buf = np.zeros(2 * 7)
    assert fftlib.h_fft(_p(buf), 7, 0, 0) == -1


@pytest.mark.parametrize("N", [4, 8, 30, 64, 200, 1024, 4096, 10000])
def test_real_fft_pre_post_steps(fftlib, N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal(N)
    X = np.zeros(2 * (N // 2 + 1))
    assert fftlib.h_rfft(_p(x), N, _p(X)) == 0
    Xc = X[0::2] + 1j * X[1::2]
    want = np.fft.rfft(x)
    assert np.max(np.abs(Xc - want)) < 1e-13 * np.sqrt(N) * np.max(np.abs(want))
    back = np.zeros(N)
    assert fftlib.h_irfft(_p(np.ascontiguousarray(X)), N, _p(back)) == 0
    assert np.max(np.abs(back / N - x)) < 1e-13 * np.sqrt(N)


def test_ln_pos_within_two_ulp(mathlib):
    rng = np.random.default_rng(0)
    x = np.concatenate([
        np.exp(rng.uniform(-700, 700, 20000)),            # whole exponent range
        1.0 + rng.uniform(-1e-3, 1e-3, 5000),             # near 1 (result near 0)
        rng.uniform(0.5, 2.0, 20000),                     # reduction boundaries
        np.array([1.0, 2.0, 0.5, np.sqrt(0.5), np.sqrt(2.0), 5e-324, 2.2250738585072014e-308,
                  1.7976931348623157e308, np.nextafter(np.sqrt(0.5), 0), np.float32(3.7)]),
    ])
    out = np.zeros_like(x)
    mathlib.h_ln(_p(x), x.size, _p(out))
    want = np.log(x)
    ulp = np.spacing(np.abs(want))
    err = np.abs(out - want)
    # two ulp of the result, or 2^-53 absolute around x = 1 where ln x -> 0
    assert np.all(err <= np.maximum(2 * ulp, 1.2e-16 * np.abs(x - 1) + 1e-300)), float(np.max(err / ulp))


def test_ln_f32_table_log(mathlib):
    """The table log the FP32 light-curve kernels use, against numpy's FP64
    log of the same float: every table interval and both of its ends, the
    whole float exponent range, subnormals, and values near 1."""
    rng = np.random.default_rng(1)
    f = np.arange(1 << 23, dtype=np.uint32)[:: 97]                      # fraction patterns
    mant = (f | np.uint32(0x3F800000)).view(np.float32)                  # [1, 2)
    ends = np.array([1.0 + i / 256.0 for i in range(257)], np.float32)
    x = np.concatenate([
        mant, ends, np.nextafter(ends, np.float32(0)),
        np.exp(rng.uniform(-103.0, 88.7, 50000)).astype(np.float32),
        (1.0 + rng.uniform(-1e-3, 1e-3, 20000)).astype(np.float32),
        np.array([1.0, 2.0, 0.5, 1e-45, 1.1754942e-38, 1.17549435e-38, 3.4028235e38], np.float32),
        np.nextafter(np.float32(1.0), np.float32(0)) * np.ones(1, np.float32),
        np.nextafter(np.float32(1.0), np.float32(2)) * np.ones(1, np.float32),
    ]).astype(np.float32)
    x = x[x > 0]
    out = np.zeros(x.size)
    mathlib.h_ln_f32(_p(x), x.size, _p(out))
    want = np.log(x.astype(np.float64))
    err = np.abs(out - want)
    assert np.all(err <= 2 * np.spacing(np.abs(want)) + 3e-16), float(np.max(err))
