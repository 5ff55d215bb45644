"""NTT parity through the C ABI: bit-exact against the CPU oracle and the
reference's LDE known answers (tests/ntt_test.cu:633-725)."""
import json
import os

import numpy as np
import pytest

from gpu_helpers import CopyHooks, ntt

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def brev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2) if bits else 0


@pytest.mark.parametrize("log_n", [1, 2, 5, 8, 11, 12, 13, 16, 19, 20])
@pytest.mark.parametrize("bitrev,inverse", [(False, False), (True, False), (False, True), (True, True)])
def test_ntt_matches_oracle(bellman, oracle, log_n, bitrev, inverse):
    B, dev = bellman
    x = oracle.random_fr(0xBC000001 + log_n, 1 << log_n)
    got = ntt(dev, x, log_n, bitrev=bitrev, inverse=inverse)
    exp = oracle.ntt(x, log_n, bitrev, inverse)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("log_n,log_ext", [(1, 1), (4, 2), (10, 3), (13, 2), (16, 2)])
@pytest.mark.parametrize("bitrev", [False, True])
def test_ntt_coset_forward_and_inverse(bellman, oracle, log_n, log_ext, bitrev):
    B, dev = bellman
    x = oracle.random_fr(77 + log_n, 1 << log_n)
    for coset in sorted({0, 1, (1 << log_ext) - 1}):
        fwd = ntt(dev, x, log_n, bitrev=bitrev, log_ext=log_ext, coset=coset)
        assert np.array_equal(fwd, oracle.ntt(x, log_n, bitrev, False, log_ext, coset))
        back = ntt(dev, fwd, log_n, bitrev=not bitrev, inverse=True, log_ext=log_ext, coset=coset)
        assert np.array_equal(back, x)


def test_ntt_lde_known_answers(bellman, oracle):
    B, dev = bellman
    kats = json.load(open(os.path.join(GOLDEN, "ntt_lde_kat.json")))
    for name, remap in (("correctness_lde_2", False), ("correctness_lde_4", True)):
        k = kats[name]
        L, e = k["log_values_count"], k["log_extension_degree"]
        inp = oracle.to_limbs([int(v, 16) for v in k["inputs"]])
        exp = [int(v, 16) for v in k["expected_outputs"]]
        for c in range(1 << e):
            out = oracle.from_limbs(ntt(dev, inp, L, log_ext=e, coset=c))
            blk = brev(c, e) if remap else c
            assert out == exp[blk << L:(blk + 1) << L]


@pytest.mark.parametrize("host_in,host_out,can_overwrite", [(True, False, False), (False, True, False), (False, True, True), (True, True, False)])
def test_ntt_host_pointers(bellman, oracle, host_in, host_out, can_overwrite):
    B, dev = bellman
    L = 14
    x = oracle.random_fr(5, 1 << L)
    got = ntt(dev, x, L, host_in=host_in, host_out=host_out, can_overwrite=can_overwrite)
    assert np.array_equal(got, oracle.ntt(x, L))


def test_ntt_2_24_round_trip_and_sampled_evaluations(bellman, oracle):
    """Full metric size: forward n2b then inverse b2n is the identity, and
    sampled outputs equal brute-force polynomial evaluations."""
    B, dev = bellman
    L = 24
    x = oracle.random_fr(0xBC000003, 1 << L)
    y = ntt(dev, x, L)
    R = oracle.R
    w = pow(pow(7, (R - 1) >> 28, R), 1 << (28 - L), R)
    coeffs = None
    for k in (0, 1, 12345, (1 << L) - 1):
        z = oracle.to_mont([pow(w, k, R)])
        assert np.array_equal(y[brev(k, L)], oracle.poly_eval(x, z)[0])
    z = ntt(dev, y, L, bitrev=True, inverse=True)
    assert np.array_equal(z, x)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("log_n", [26, 28])
def test_ntt_maximum_sizes_round_trip(bellman, oracle, log_n):
    """The largest transforms the API takes (2^28 elements, 8 GiB, as in the
    reference's benchmark range, tests/ntt_test.cu:727-736): forward natural ->
    bit-reversed then inverse bit-reversed -> natural on the device is the
    identity, and output 0 is P(1) = sum x_i (Montgomery form is linear, so the
    limb-wise sum of the stored inputs mod r is the stored output)."""
    B, dev = bellman
    n = 1 << log_n
    rng = np.random.default_rng(0xBC000003 + log_n)
    x = rng.integers(0, 1 << 32, size=(n, 8), dtype=np.uint32)
    x[:, 7] %= 0x30644E72  # below r's top word: every value is below r
    total = sum(int(x[:, i].sum(dtype=np.uint64)) << (32 * i) for i in range(8)) % oracle.R
    d_x = dev.upload(x)
    d_y = dev.malloc(32 * n)
    try:
        for inp, out, bitrev, inverse in ((d_x, d_y, False, False), (d_y, d_x, True, True)):
            cfg = B.ntt_configuration()
            cfg.mem_pool, cfg.stream, cfg.inputs, cfg.outputs = dev.pool, dev.stream, inp, out
            cfg.log_values_count, cfg.bit_reversed_inputs, cfg.inverse = log_n, bitrev, inverse
            B.call("ntt_execute_async", cfg)
            if not inverse:
                dev.sync()
                assert oracle.from_limbs(dev.download(d_y, (1, 8)))[0] == total
        dev.sync()
        assert np.array_equal(dev.download(d_x, (n, 8)), x)
    finally:
        dev.free(d_x)
        dev.free(d_y)


def test_lde_assembled_in_whole_domain_bit_reversed_order(bellman, oracle):
    """Prover-side assembly of a 4x LDE at 2^16: running coset brev_2(c) for
    output block c gives, at position j of the 2^18 outputs, P(g w^brev(j)),
    i.e. the size-2^18 natural -> bit-reversed transform of (g^i x_i) padded
    with zeros (the ordering the reference's lde_4 answers follow,
    tests/ntt_test.cu:633-725)."""
    B, dev = bellman
    L, e = 16, 2
    n = 1 << L
    x = oracle.random_fr(0xBC000060, n)
    got = np.concatenate([ntt(dev, x, L, log_ext=e, coset=brev(c, e)) for c in range(1 << e)])
    y = np.zeros(((1 << e) * n, 8), dtype=np.uint32)
    y[:n] = oracle.multiply_by_powers(x, oracle.to_mont([7]))
    assert np.array_equal(got, oracle.ntt(y, L + e))


@pytest.mark.parametrize("host_in,host_out", [(True, True), (True, False), (False, True), (False, False)])
def test_ntt_copy_hooks(bellman, oracle, host_in, host_out):
    """The h2d/d2h events and callbacks of ntt_configuration
    (src/ntt.cu:170-180,220-229): the h2d callback runs once after host inputs
    are copied in, the d2h callback once after host outputs are copied out, in
    that order and with their user data, and the events complete; a buffer
    already on the device is not copied and its callback does not run.  The
    outputs are the oracle's."""
    B, dev = bellman
    L = 10
    x = oracle.random_fr(0xBC000073, 1 << L)
    hooks = CopyHooks()
    out_host = np.zeros((1 << L, 8), dtype=np.uint32)
    d_in = None if host_in else dev.upload(x)
    d_out = None if host_out else dev.malloc(32 << L)
    try:
        cfg = B.ntt_configuration()
        cfg.mem_pool, cfg.stream = dev.pool, dev.stream
        cfg.inputs = x.ctypes.data if host_in else d_in
        cfg.outputs = out_host.ctypes.data if host_out else d_out
        cfg.log_values_count = L
        hooks.attach(cfg)
        B.call("ntt_execute_async", cfg)
        dev.sync()
        got = out_host if host_out else dev.download(d_out, (1 << L, 8))
        assert np.array_equal(got, oracle.ntt(x, L, False, False))
        expected = [CopyHooks.H2D] * host_in + [CopyHooks.D2H] * host_out
        assert hooks.wait_calls(len(expected)) == expected
        assert hooks.events_complete()
    finally:
        hooks.close()
        for p in (d_in, d_out):
            if p is not None:
                dev.free(p)
