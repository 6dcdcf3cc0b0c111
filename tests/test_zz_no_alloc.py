"""lqcs_state_expectation_terms_noalloc (the allocation-free variant of
lqcs_state_expectation_terms) through the C-ABI: on two calls in a row, the
second on a smaller register with fewer terms, its value equals the
allocating entry point's within 1e-12 (the same partial-sum kernels; the
final weighted sum runs on the device instead of the host, which may round
its last bit differently), and a term count beyond LQCS_NOALLOC_MAX_TERMS is
refused with INVALID_ARGUMENT."""
import ctypes
import os

import numpy as np
import pytest

from litequantumcomputingsimulator_amd import LIB_DIR

pytestmark = pytest.mark.gpu

MAX_TERMS = 512
u64p = ctypes.POINTER(ctypes.c_uint64)
u32p = ctypes.POINTER(ctypes.c_uint32)
dp = ctypes.POINTER(ctypes.c_double)


def _lib():
    lib = ctypes.CDLL(os.path.join(LIB_DIR, "liblqcs_device.so"))
    lib.lqcs_dev_last_error.restype = ctypes.c_char_p
    return lib


def _random_register(lib, n, seed):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    a /= np.linalg.norm(a)
    st = ctypes.c_void_p()
    assert lib.lqcs_state_create(ctypes.c_uint32(n), ctypes.byref(st)) == 0
    buf = np.ascontiguousarray(np.stack([a.real, a.imag], axis=1).reshape(-1))
    assert lib.lqcs_state_upload(st, buf.ctypes.data_as(dp)) == 0
    return st


def _random_terms(n, count, seed):
    rng = np.random.default_rng(seed)
    xs, zs, ys = [], [], []
    while len(xs) < count:
        letters = rng.integers(0, 4, size=n)  # I X Y Z
        if not letters.any():
            continue
        x = zy = ny = 0
        for q, k in enumerate(letters):
            if k in (1, 2):
                x |= 1 << q
            if k in (2, 3):
                zy |= 1 << q
            ny += int(k == 2)
        xs.append(x)
        zs.append(zy)
        ys.append(ny)
    return (np.array(xs, np.uint64), np.array(zs, np.uint64), np.array(ys, np.uint32),
            rng.normal(size=count).astype(np.float64))


def _both(lib, st, terms):
    xs, zs, ys, cs = terms
    want, got = ctypes.c_double(), ctypes.c_double()
    args = (xs.ctypes.data_as(u64p), zs.ctypes.data_as(u64p), ys.ctypes.data_as(u32p), cs.ctypes.data_as(dp),
            ctypes.c_uint64(len(cs)))
    assert lib.lqcs_state_expectation_terms(st, *args, ctypes.byref(want)) == 0
    rc = lib.lqcs_state_expectation_terms_noalloc(st, *args, ctypes.byref(got))
    assert rc == 0, lib.lqcs_dev_last_error()
    return want.value, got.value


def test_noalloc_expectation_equals_allocating_entry_point():
    lib = _lib()
    assert lib.lqcs_dev_init(0) == 0
    big = _random_register(lib, 16, seed=1)
    small = _random_register(lib, 9, seed=2)
    try:
        for st, n, count, seed in ((big, 16, 300, 3), (small, 9, 37, 4)):  # second call: smaller input
            want, got = _both(lib, st, _random_terms(n, count, seed))
            assert abs(got - want) <= 1e-12, (n, got, want)
            assert abs(want) > 1e-6  # not vacuous
        # the maximum is accepted, one more is refused before any device work
        want, got = _both(lib, small, _random_terms(9, MAX_TERMS, 5))
        assert abs(got - want) <= 1e-12
        xs, zs, ys, cs = _random_terms(9, MAX_TERMS + 1, 6)
        out = ctypes.c_double(7.0)
        rc = lib.lqcs_state_expectation_terms_noalloc(small, xs.ctypes.data_as(u64p), zs.ctypes.data_as(u64p),
                                                      ys.ctypes.data_as(u32p), cs.ctypes.data_as(dp),
                                                      ctypes.c_uint64(MAX_TERMS + 1), ctypes.byref(out))
        assert rc == 1 and b"at most 512" in lib.lqcs_dev_last_error()
        # zero terms: 0, like the allocating entry point
        assert lib.lqcs_state_expectation_terms_noalloc(small, None, None, None, None, ctypes.c_uint64(0),
                                                        ctypes.byref(out)) == 0 and out.value == 0.0
    finally:
        lib.lqcs_state_destroy(big)
        lib.lqcs_state_destroy(small)
