"""CPU check of the LDS FFT stages (curig_amd/csrc/fft_core.h): the same code the
GPU kernels run, with threads emulated by loops, against numpy.fft."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "fft_core_check.cpp")
CSRC = os.path.join(ROOT, "curig_amd", "csrc")


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("fftc") / "libfftc.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", CSRC, "-o", out, SRC])
    L = ctypes.cdll.LoadLibrary(out)
    L.fftc_check_run.argtypes = [ctypes.c_int] * 7 + [np.ctypeslib.ndpointer(np.float64, flags="C"), ctypes.c_int,
                                                      ctypes.c_int, ctypes.c_int]
    L.fftc_check_run.restype = ctypes.c_int
    return L


def run(lib, x, n, count, seq_stride, elem_stride, seq_fast, sign, dbl, n_tw=None, E=16, maxr=16):
    data = np.ascontiguousarray(np.stack([x.real, x.imag], -1).reshape(-1))
    assert lib.fftc_check_run(n, count, seq_stride, elem_stride, seq_fast, sign, dbl, data, n_tw or n, E, maxr) == 0
    return data.reshape(-1, 2) @ np.array([1, 1j])


@pytest.mark.parametrize("n", [16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384])
@pytest.mark.parametrize("sign", [1, -1])
def test_single_sequence(lib, n, sign):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    got = run(lib, x, n, 1, n, 1, 0, sign, 1)
    want = np.fft.ifft(x) * n if sign > 0 else np.fft.fft(x)
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))
    got32 = run(lib, x, n, 1, n, 1, 0, sign, 0)
    assert np.sqrt(np.mean(np.abs(got32 - want) ** 2)) < 3e-7 * np.sqrt(np.mean(np.abs(want) ** 2))


@pytest.mark.parametrize("n,count", [(16, 32), (64, 16), (128, 32), (128, 16), (256, 8)])
def test_batched_columns(lib, n, count):
    # sequences interleaved as columns of an [n][count] block: elem_stride = count, seq_stride = 1
    rng = np.random.default_rng(n + count)
    x = rng.standard_normal((count, n)) + 1j * rng.standard_normal((count, n))
    got = run(lib, x.reshape(-1), n, count, 1, count, 1, -1, 1).reshape(count, n)
    np.testing.assert_allclose(got, np.fft.fft(x, axis=1), atol=1e-11 * np.abs(x).sum() / count)


def test_twiddle_table_multiple_of_n(lib):
    # a table for a longer transform (n_tw = 4 n) gives the same result
    n = 256
    rng = np.random.default_rng(1)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    got = run(lib, x, n, 1, n, 1, 0, 1, 1, n_tw=4 * n)
    np.testing.assert_allclose(got, np.fft.ifft(x) * n, atol=1e-10)


@pytest.mark.parametrize("n,count", [(64, 64), (256, 16), (1024, 4), (4096, 1)])
def test_row_batches(lib, n, count):
    # several contiguous rows per workgroup: seq_stride = n, elem_stride = 1
    rng = np.random.default_rng(n * count)
    x = rng.standard_normal((count, n)) + 1j * rng.standard_normal((count, n))
    got = run(lib, x.reshape(-1), n, count, n, 1, 0, 1, 1).reshape(count, n)
    np.testing.assert_allclose(got, np.fft.ifft(x, axis=1) * n, atol=1e-11 * np.abs(x).sum() / count)


@pytest.mark.parametrize("n", [32, 256, 4096, 16384])
def test_32_elements_per_thread(lib, n):
    rng = np.random.default_rng(n + 5)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    got = run(lib, x, n, 1, n, 1, 0, 1, 1, E=32)
    want = np.fft.ifft(x) * n
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))
    got32 = run(lib, x, n, 1, n, 1, 0, 1, 0, E=32)
    assert np.sqrt(np.mean(np.abs(got32 - want) ** 2)) < 3e-7 * np.sqrt(np.mean(np.abs(want) ** 2))


@pytest.mark.parametrize("n,count", [(16, 1), (64, 64), (512, 8), (4096, 1), (16384, 1)])
@pytest.mark.parametrize("sign", [1, -1])
def test_radix8_stages(lib, n, count, sign):
    # radix-8 stages (the row kernels' register-lean variant)
    rng = np.random.default_rng(n + count + sign)
    x = rng.standard_normal((count, n)) + 1j * rng.standard_normal((count, n))
    got = run(lib, x.reshape(-1), n, count, n, 1, 0, sign, 1, maxr=8).reshape(count, n)
    want = np.fft.ifft(x, axis=1) * n if sign > 0 else np.fft.fft(x, axis=1)
    np.testing.assert_allclose(got, want, atol=1e-11 * np.abs(x).sum() / count)
    got32 = run(lib, x.reshape(-1), n, count, n, 1, 0, sign, 0, maxr=8).reshape(count, n)
    assert np.sqrt(np.mean(np.abs(got32 - want) ** 2)) < 3e-7 * np.sqrt(np.mean(np.abs(want) ** 2))


@pytest.mark.parametrize("n", [64, 256, 1024, 4096, 8192, 16384])
@pytest.mark.parametrize("sign", [1, -1])
def test_two_level_twiddles(lib, n, sign):
    """The kernels' LDS twiddles: products of two sqrt(n)-entry tables."""
    lib.fftc_check_run2.argtypes = [ctypes.c_int] * 7 + [np.ctypeslib.ndpointer(np.float64, flags="C"),
                                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.fftc_check_run2.restype = ctypes.c_int
    rng = np.random.default_rng(n + 1)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    want = np.fft.ifft(x) * n if sign > 0 else np.fft.fft(x)
    for dbl, tol in ((1, 1e-12), (0, 3e-7)):
        for maxr in (8, 16):
            data = np.ascontiguousarray(np.stack([x.real, x.imag], -1).reshape(-1))
            assert lib.fftc_check_run2(n, 1, n, 1, 0, sign, dbl, data, n, 16, maxr, 1) == 0
            got = data.reshape(-1, 2) @ np.array([1, 1j])
            err = np.sqrt(np.mean(np.abs(got - want) ** 2)) / np.sqrt(np.mean(np.abs(want) ** 2))
            assert err < tol, (dbl, maxr, err)
