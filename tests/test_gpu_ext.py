"""Extension entry points through the C ABI: fixed-base generator multiples
against the oracle's [k]G, including the comb's byte boundaries and zero."""
import ctypes

import numpy as np
import pytest

import bellman_amd as B

pytestmark = pytest.mark.gpu


def test_generator_multiples_match_oracle(bellman, oracle):
    _, dev = bellman
    R = oracle.R
    ks = [1, 2, 3, 255, 256, 257, 65535, 1 << 64, R - 1, R - 2, (1 << 253) + 12345]
    ks += [int.from_bytes(bytes((i * 37 + j * 11) & 0xFF for j in range(32)), "little") % R for i in range(53)]
    d_s = dev.upload(oracle.to_limbs(ks))
    d_p = dev.malloc(64 * len(ks))
    B.call("bc_amd_generator_multiples", ctypes.c_void_p(d_s), ctypes.c_void_p(d_p), ctypes.c_uint(len(ks)), dev.stream)
    dev.sync()
    got = dev.download(d_p, (len(ks), 16))
    for i, k in enumerate(ks):
        assert np.array_equal(got[i], oracle.mul_generator(k)), (i, k)
    dev.free(d_s)
    dev.free(d_p)


def test_generator_multiples_of_zero_are_zero_encoded(bellman, oracle):
    _, dev = bellman
    d_s = dev.upload(oracle.to_limbs([0, 0]))
    d_p = dev.malloc(128)
    B.call("bc_amd_generator_multiples", ctypes.c_void_p(d_s), ctypes.c_void_p(d_p), ctypes.c_uint(2), dev.stream)
    dev.sync()
    assert not dev.download(d_p, (2, 16)).any()
    dev.free(d_s)
    dev.free(d_p)
