"""NTT -> MSM chains through the C ABI, the prover flows of the reference's
tests/bc_test.cu:69-170: a stream-ordered temporary from
bc_malloc_from_pool_async holds the NTT's input, the NTT runs in place, and its
output is used directly as MSM scalars (whole, or as four slices in series on
one stream, or in parallel on four streams that wait on an NTT event and are
destroyed right after their work is enqueued); the temporary is released with
bc_free_async and the MSM results land in host memory.  Each MSM fold is
checked against the closed form [sum k_i s_i]G with k the oracle's NTT output
and bases [s0 + i d]G."""
import ctypes

import numpy as np
import pytest

import bellman_amd as B

pytestmark = pytest.mark.gpu

S0, D = 0xBC000081, 0x6A09E667F3BCC908


@pytest.fixture(scope="module")
def chain_bases(bellman, oracle):
    _, dev = bellman
    n = 1 << 12
    db = dev.upload(oracle.gen_bases_progression(S0, D, n))
    yield db, oracle.fr_progression(S0, D, n)
    dev.free(db)


def new_stream(blocking=True):
    st = B.bc_stream()
    B.call("bc_stream_create", ctypes.byref(st), ctypes.c_bool(blocking))
    return st


def new_event():
    ev = B.bc_event()
    B.call("bc_event_create", ctypes.byref(ev), ctypes.c_bool(True), ctypes.c_bool(True))
    return ev


def pool_temp_with(dev, stream, host):
    """bc_malloc_from_pool_async + bc_memcpy_async of a host array."""
    p = ctypes.c_void_p()
    B.call("bc_malloc_from_pool_async", ctypes.byref(p), ctypes.c_size_t(host.nbytes), dev.pool, stream)
    B.call("bc_memcpy_async", p, host.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(host.nbytes), stream)
    return p.value


def ntt_in_place(dev, stream, d_values, L):
    cfg = B.ntt_configuration()
    cfg.mem_pool, cfg.stream, cfg.inputs, cfg.outputs, cfg.log_values_count = dev.pool, stream, d_values, d_values, L
    B.call("ntt_execute_async", cfg)


def msm_to_host(dev, stream, d_bases, d_scalars, results, L):
    cfg = B.msm_configuration()
    cfg.mem_pool, cfg.stream = dev.pool, stream
    cfg.bases, cfg.scalars, cfg.results, cfg.log_scalars_count = d_bases, d_scalars, results.ctypes.data, L
    B.call("msm_execute_async", cfg)


def expected_fold(oracle, s_mont, k_mont):
    t = oracle.from_mont(oracle.fr_dot(k_mont, s_mont[: k_mont.shape[0]]))[0]
    return oracle.mul_generator(t)


def check_fold(oracle, res, s_mont, k):
    assert np.array_equal(oracle.to_affine(oracle.msm_fold(res))[0], expected_fold(oracle, s_mont, k))


def test_ntt_into_msm(bellman, oracle, chain_bases):
    """tests/bc_test.cu:69-93 at 2^12: NTT in place on a pool temporary, then
    an MSM over its output, results to the host."""
    _, dev = bellman
    db, s_mont = chain_bases
    L = 12
    x = oracle.random_fr(0xBC000082, 1 << L)
    results = np.zeros((254, 24), dtype=np.uint32)
    st = new_stream()
    try:
        d_values = pool_temp_with(dev, st, x)
        ntt_in_place(dev, st, d_values, L)
        msm_to_host(dev, st, db, d_values, results, L)
        B.call("bc_free_async", ctypes.c_void_p(d_values), st)
        B.call("bc_stream_synchronize", st)
    finally:
        B.call("bc_stream_destroy", st)
    check_fold(oracle, results, s_mont, oracle.ntt(x, L))


def test_ntt_into_four_msms_in_series(bellman, oracle, chain_bases):
    """tests/bc_test.cu:95-125: a 2^12 NTT, then four 2^10 MSMs over its
    quarters, one after another on the same stream."""
    _, dev = bellman
    db, s_mont = chain_bases
    L, Lm = 12, 10
    x = oracle.random_fr(0xBC000083, 1 << L)
    results = [np.zeros((254, 24), dtype=np.uint32) for _ in range(4)]
    st = new_stream()
    try:
        d_values = pool_temp_with(dev, st, x)
        ntt_in_place(dev, st, d_values, L)
        for i in range(4):
            msm_to_host(dev, st, db, d_values + 32 * (i << Lm), results[i], Lm)
        B.call("bc_free_async", ctypes.c_void_p(d_values), st)
        B.call("bc_stream_synchronize", st)
    finally:
        B.call("bc_stream_destroy", st)
    y = oracle.ntt(x, L)
    for i in range(4):
        check_fold(oracle, results[i], s_mont, y[i << Lm:(i + 1) << Lm])


def test_ntt_into_four_msms_in_parallel(bellman, oracle, chain_bases):
    """tests/bc_test.cu:127-170: after the NTT an event is recorded; four MSMs
    run on four streams that wait on it, each stream records an event the main
    stream waits on, and each stream and event is destroyed right after its
    work is enqueued; the main stream then frees the temporary."""
    _, dev = bellman
    db, s_mont = chain_bases
    L, Lm = 12, 10
    x = oracle.random_fr(0xBC000084, 1 << L)
    results = [np.zeros((254, 24), dtype=np.uint32) for _ in range(4)]
    st = new_stream()
    try:
        d_values = pool_temp_with(dev, st, x)
        ntt_in_place(dev, st, d_values, L)
        ntt_done = new_event()
        B.call("bc_event_record", ntt_done, st)
        for i in range(4):
            ms = new_stream()
            B.call("bc_stream_wait_event", ms, ntt_done)
            msm_to_host(dev, ms, db, d_values + 32 * (i << Lm), results[i], Lm)
            msm_done = new_event()
            B.call("bc_event_record", msm_done, ms)
            B.call("bc_stream_wait_event", st, msm_done)
            B.call("bc_event_destroy", msm_done)
            B.call("bc_stream_destroy", ms)
        B.call("bc_free_async", ctypes.c_void_p(d_values), st)
        B.call("bc_stream_synchronize", st)
        B.call("bc_event_destroy", ntt_done)
    finally:
        B.call("bc_stream_destroy", st)
    y = oracle.ntt(x, L)
    for i in range(4):
        check_fold(oracle, results[i], s_mont, y[i << Lm:(i + 1) << Lm])


def test_lde_then_msm_over_each_coset(bellman, oracle, chain_bases):
    """A prover's commitment step: a 4x coset extension of 2^10 coefficients
    (four coset NTTs into slices of one pool temporary), then one MSM of 2^12
    over the whole extended vector: the fold equals the closed form with the
    oracle's four coset outputs as scalars."""
    _, dev = bellman
    db, s_mont = chain_bases
    L, e = 10, 2
    x = oracle.random_fr(0xBC000085, 1 << L)
    results = np.zeros((254, 24), dtype=np.uint32)
    st = new_stream()
    try:
        d_x = pool_temp_with(dev, st, x)
        ext = ctypes.c_void_p()
        B.call("bc_malloc_from_pool_async", ctypes.byref(ext), ctypes.c_size_t(32 << (L + e)), dev.pool, st)
        for c in range(1 << e):
            cfg = B.ntt_configuration()
            cfg.mem_pool, cfg.stream, cfg.inputs, cfg.outputs = dev.pool, st, d_x, ext.value + 32 * (c << L)
            cfg.log_values_count, cfg.log_extension_degree, cfg.coset_index = L, e, c
            B.call("ntt_execute_async", cfg)
        msm_to_host(dev, st, db, ext.value, results, L + e)
        B.call("bc_free_async", ext, st)
        B.call("bc_free_async", ctypes.c_void_p(d_x), st)
        B.call("bc_stream_synchronize", st)
    finally:
        B.call("bc_stream_destroy", st)
    k = np.concatenate([oracle.ntt(x, L, False, False, e, c) for c in range(1 << e)])
    check_fold(oracle, results, s_mont, k)
