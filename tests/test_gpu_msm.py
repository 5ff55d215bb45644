"""MSM parity through the C ABI.  The 254 per-bit results are compared with
the CPU oracle projectively (after affine normalisation: the reference's raw
Jacobian limbs are not canonical), and their fold against the closed form
[sum k_i s_i]G for bases P_i = [s_i]G (tests/msm_test.cu:107-141)."""
import numpy as np
import pytest

from gpu_helpers import CopyHooks, msm, msm_batch

pytestmark = pytest.mark.gpu


def closed_form(oracle, s_ints, k_mont):
    ks = oracle.from_mont(k_mont)
    return oracle.mul_generator(sum(a * b for a, b in zip(ks, s_ints)) % oracle.R)


@pytest.mark.parametrize("log_n", [0, 1, 4, 7, 10])
def test_msm_per_bit_results_match_oracle(bellman, oracle, log_n):
    B, dev = bellman
    n = 1 << log_n
    s = oracle.from_limbs(oracle.random_fr(0xBC000020 + log_n, n, montgomery=False))
    bases = oracle.gen_bases(s)
    k = oracle.msm_scalars(0xBC000021 + log_n, n)
    db = dev.upload(bases)
    got = msm(dev, db, k, log_n)
    dev.free(db)
    exp = oracle.msm(bases, k, log_n)
    assert np.array_equal(oracle.to_affine(got), oracle.to_affine(exp))
    assert np.array_equal(oracle.to_affine(oracle.msm_fold(got))[0], closed_form(oracle, s, k))


@pytest.mark.parametrize("log_n", [12, 14, 16])
def test_msm_fold_closed_form(bellman, oracle, log_n):
    B, dev = bellman
    n = 1 << log_n
    s0, d = 0x1234567 + log_n, 0xABCDEF1
    bases = oracle.gen_bases_progression(s0, d, n)
    k = oracle.msm_scalars(0xBC000022 + log_n, n)
    db = dev.upload(bases)
    got = msm(dev, db, k, log_n)
    dev.free(db)
    s = [(s0 + i * d) % oracle.R for i in range(n)]
    assert np.array_equal(oracle.to_affine(oracle.msm_fold(got))[0], closed_form(oracle, s, k))


@pytest.mark.parametrize("host_scalars,host_results", [(True, False), (False, True), (True, True)])
def test_msm_host_pointers(bellman, oracle, host_scalars, host_results):
    B, dev = bellman
    L = 9
    n = 1 << L
    s = oracle.from_limbs(oracle.random_fr(31, n, montgomery=False))
    bases = oracle.gen_bases(s)
    k = oracle.msm_scalars(32, n)
    db = dev.upload(bases)
    got = msm(dev, db, k, L, host_scalars=host_scalars, host_results=host_results)
    dev.free(db)
    assert np.array_equal(oracle.to_affine(oracle.msm_fold(got))[0], closed_form(oracle, s, k))


def test_msm_edge_scalars(bellman, oracle):
    """All-zero, all-one, r-1 and repeated bases (doubling inside a bucket)."""
    B, dev = bellman
    L = 6
    n = 1 << L
    R = oracle.R
    s = [5] * (n // 2) + [(7 + i) for i in range(n // 2)]  # repeated base P = [5]G
    bases = oracle.gen_bases(s)
    db = dev.upload(bases)
    for ks in ([0] * n, [1] * n, [R - 1] * n, [(i * 0x9E3779B97F4A7C15) % R for i in range(n)]):
        k = oracle.to_mont(ks)
        got = msm(dev, db, k, L)
        exp = closed_form(oracle, s, k)
        assert np.array_equal(oracle.to_affine(oracle.msm_fold(got))[0], exp)
        assert np.array_equal(oracle.to_affine(got), oracle.to_affine(oracle.msm(bases, k, L)))
    dev.free(db)


def test_msm_cancelling_and_repeated_points_in_one_bucket(bellman, oracle):
    """Equal bases whose scalars are k and r - k meet in the same bucket with
    opposite signs (the running sum becomes infinity and starts again), equal
    bases with equal scalars double, and a bucket meets a point, its negation
    and the point again; the per-bit results equal the oracle's and the fold
    its closed form."""
    import random

    B, dev = bellman
    L = 8
    n = 1 << L
    R = oracle.R
    rng = random.Random(0xBC0CA9CE)
    s = [rng.randrange(1, R) for _ in range(n)]
    ks = [rng.randrange(R) for _ in range(n)]
    k0 = rng.randrange(1, R)
    s[1] = s[2] = s[0]
    ks[0], ks[1], ks[2] = k0, R - k0, k0  # +P, -P, +P in each window's bucket
    s[4] = s[3]
    ks[4] = R - ks[3]  # cancelling pair
    s[6] = s[5]
    ks[6] = ks[5]  # doubling
    for j in range(16):  # more cancelling pairs, in other buckets and chunks
        a = 7 + 2 * j
        s[a + 1] = s[a]
        ks[a + 1] = (R - ks[a]) % R
    bases = oracle.gen_bases(s)
    k = oracle.to_mont(ks)
    db = dev.upload(bases)
    got = msm(dev, db, k, L)
    assert np.array_equal(oracle.to_affine(got), oracle.to_affine(oracle.msm(bases, k, L)))
    assert np.array_equal(oracle.to_affine(oracle.msm_fold(got))[0], closed_form(oracle, s, k))
    # a batch job with long chunks (32 entries, window 6: a bucket's entries
    # mostly in one chunk, so one thread meets +P, -P, +P in a row)
    ds, dr = dev.upload(k), dev.malloc(254 * 96)
    msm_batch(dev, [B.msm_job(db, ds, dr, L, 6, 32)])
    got = dev.download(dr, (254, 24))
    assert np.array_equal(oracle.to_affine(got), oracle.to_affine(oracle.msm(bases, k, L, 6)))
    for p in (ds, dr, db):
        dev.free(p)


def test_msm_config_2_20(bellman, oracle):
    """The 2^20 MSM workload: n = 2^20, scalars 70% uniform / 20% < 2^64 / 10% in {0,1}."""
    B, dev = bellman
    L = 20
    n = 1 << L
    s0, d = 0xBC000002, 0x1F2E3D4C5B6A7988
    bases = oracle.gen_bases_progression(s0, d, n)
    k = oracle.msm_scalars(0xBC000002, n)
    db = dev.upload(bases)
    got = msm(dev, db, k, L)
    dev.free(db)
    ks = oracle.from_mont(k)
    R = oracle.R
    t = (s0 * sum(ks) + d * sum(i * kv for i, kv in enumerate(ks))) % R
    assert np.array_equal(oracle.to_affine(oracle.msm_fold(got))[0], oracle.mul_generator(t))


def test_msms_overlap_on_four_streams(bellman, oracle):
    """Four plain msm_execute_async calls in flight on four streams at once (as
    tests/bc_test.cu:126-170 overlaps them); each result equals the same MSM
    run alone, byte for byte, and folds to its closed form."""
    import ctypes

    B, dev = bellman
    s0, d = 0xBC000061, 0x2F3E4D5C6B7A8999
    bases = oracle.gen_bases_progression(s0, d, 1 << 14)
    db = dev.upload(bases)
    logs = [(s0 + i * d) % oracle.R for i in range(1 << 14)]
    sizes = [14, 12, 13, 10]
    scal = [oracle.msm_scalars(0xBC000062 + j, 1 << L) for j, L in enumerate(sizes)]
    alone = [msm(dev, db, k, L) for k, L in zip(scal, sizes)]
    streams, d_s, d_r = [], [], []
    try:
        for j, L in enumerate(sizes):
            st = B.bc_stream()
            B.call("bc_stream_create", ctypes.byref(st), ctypes.c_bool(False))
            streams.append(st)
            d_s.append(dev.upload(scal[j]))
            d_r.append(dev.malloc(254 * 96))
        for j, L in enumerate(sizes):
            cfg = B.msm_configuration()
            cfg.mem_pool, cfg.stream = dev.pool, streams[j]
            cfg.bases, cfg.scalars, cfg.results = db, d_s[j], d_r[j]
            cfg.log_scalars_count = L
            B.call("msm_execute_async", cfg)
        for st in streams:
            B.call("bc_stream_synchronize", st)
        for j, L in enumerate(sizes):
            got = dev.download(d_r[j], (254, 24))
            assert np.array_equal(got, alone[j]), L
            assert np.array_equal(oracle.to_affine(oracle.msm_fold(got))[0], closed_form(oracle, logs[: 1 << L], scal[j]))
    finally:
        for st in streams:
            B.call("bc_stream_destroy", st)
        for p in d_s + d_r + [db]:
            dev.free(p)


@pytest.mark.parametrize("host_scalars,host_results", [(True, True), (True, False), (False, True), (False, False)])
def test_msm_copy_hooks(bellman, oracle, host_scalars, host_results):
    """The h2d/d2h events and callbacks of msm_configuration
    (src/msm.cu:183-191,394-400): the h2d callback runs once after host
    scalars are copied in, the d2h callback once after host results are copied
    out, in that order and with their user data, and the events complete; a
    buffer already on the device is not copied and its callback does not run.
    The results equal those of the device-buffer call."""
    B, dev = bellman
    L = 9
    s = oracle.from_limbs(oracle.random_fr(0xBC000071, 1 << L, montgomery=False))
    db = dev.upload(oracle.gen_bases(s))
    k = oracle.msm_scalars(0xBC000072, 1 << L)
    want = msm(dev, db, k, L)
    hooks = CopyHooks()
    res_host = np.zeros((254, 24), dtype=np.uint32)
    d_s = None if host_scalars else dev.upload(k)
    d_r = None if host_results else dev.malloc(254 * 96)
    try:
        cfg = B.msm_configuration()
        cfg.mem_pool, cfg.stream, cfg.bases = dev.pool, dev.stream, db
        cfg.scalars = k.ctypes.data if host_scalars else d_s
        cfg.results = res_host.ctypes.data if host_results else d_r
        cfg.log_scalars_count = L
        hooks.attach(cfg)
        B.call("msm_execute_async", cfg)
        dev.sync()
        got = res_host if host_results else dev.download(d_r, (254, 24))
        assert np.array_equal(got, want)
        expected = [CopyHooks.H2D] * host_scalars + [CopyHooks.D2H] * host_results
        assert hooks.wait_calls(len(expected)) == expected
        assert hooks.events_complete()
    finally:
        hooks.close()
        for p in (d_s, d_r, db):
            if p is not None:
                dev.free(p)


@pytest.mark.timeout(300)
def test_msm_2_26_linearity(bellman, oracle):
    """The reference's largest benchmarked MSM (2^26 points, window 22,
    tests/msm_test.cu:253-260): bases [s_i]G made on the device from random
    s_i, two random scalar vectors k1 and k2 and k = k1 + k2 (Montgomery form
    is linear).  The folded results satisfy MSM(k) = MSM(k1) + MSM(k2), the
    sum taken by the oracle as the MSM of the two folded points with scalars
    1 and 1."""
    import ctypes

    B, dev = bellman
    L = 26
    n = 1 << L
    rng = np.random.default_rng(0xBC000026)

    def below_r():
        x = rng.integers(0, 1 << 32, size=(n, 8), dtype=np.uint32)
        x[:, 7] %= 0x30644E72  # below r's top word: every value is below r
        return x

    bufs = []
    try:
        d_s = dev.upload(below_r())
        d_b = dev.malloc(64 * n)
        bufs += [d_s, d_b]
        B.call("bc_amd_generator_multiples", ctypes.c_void_p(d_s), ctypes.c_void_p(d_b), ctypes.c_uint(n), dev.stream)
        d_k1, d_k2 = dev.upload(below_r()), dev.upload(below_r())
        d_k = dev.malloc(32 * n)
        bufs += [d_k1, d_k2, d_k]
        B.call("ff_x_plus_y", ctypes.c_void_p(d_k1), ctypes.c_void_p(d_k2), ctypes.c_void_p(d_k), ctypes.c_uint(n), dev.stream)
        res = [dev.malloc(254 * 96) for _ in range(3)]
        bufs += res
        for dk, dr in zip((d_k1, d_k2, d_k), res):
            cfg = B.msm_configuration()
            cfg.mem_pool, cfg.stream = dev.pool, dev.stream
            cfg.bases, cfg.scalars, cfg.results = d_b, dk, dr
            cfg.log_scalars_count = L
            B.call("msm_execute_async", cfg)
        d_p = dev.malloc(64 * 3)
        bufs.append(d_p)
        B.call("bc_amd_msm_fold", (ctypes.c_void_p * 3)(*res), ctypes.c_void_p(d_p), ctypes.c_uint(3), dev.stream)
        dev.sync()
        got = dev.download(d_p, (3, 16))
        assert got[0].any() and got[1].any() and not np.array_equal(got[0], got[1])
        both = oracle.msm(np.ascontiguousarray(got[:2]), oracle.to_mont([1, 1]), 1)
        assert np.array_equal(oracle.to_affine(oracle.msm_fold(both))[0], got[2])
    finally:
        for p in bufs:
            dev.free(p)


def _host_buffer(B, arr, pinned):
    """A copy of arr in pageable (numpy) or pinned (bc_malloc_host) host memory; (address, keep-alive, release)."""
    import ctypes

    a = np.ascontiguousarray(arr)
    if not pinned:
        return a.ctypes.data, a, lambda: None
    p = ctypes.c_void_p()
    B.call("bc_malloc_host", ctypes.byref(p), ctypes.c_size_t(a.nbytes))
    ctypes.memmove(p.value, a.ctypes.data, a.nbytes)
    return p.value, None, lambda: B.call("bc_free_host", p)


@pytest.mark.timeout(300)
def test_msm_host_scalars_chunks_add_listed_buckets(bellman, oracle):
    """Host scalars at 2^23: four chunks of 2^21 points accumulate into the same
    windows at the full size's width (c = 19).  Each chunk's buckets hold about
    8 entries against accumulation chunks of 32, so the gather lists its
    buckets of several partial sums in LDS and adds them, with the single-sum
    ones, to what the earlier chunks left in the windows.  The per-bit results
    equal the device-input call's as points and the fold is the closed form."""
    import ctypes

    B, dev = bellman
    L = 23
    n = 1 << L
    rng = np.random.default_rng(0xBC000023)
    s = rng.integers(0, 1 << 32, size=(n, 8), dtype=np.uint32)
    s[:, 7] %= 0x30644E72  # below r's top word: every value is below r
    k = oracle.msm_scalars(0xBC000024, n)
    bufs = []
    try:
        d_s = dev.upload(s)
        d_b = dev.malloc(64 * n)
        bufs += [d_s, d_b]
        B.call("bc_amd_generator_multiples", ctypes.c_void_p(d_s), ctypes.c_void_p(d_b), ctypes.c_uint(n), dev.stream)
        r2 = oracle.to_limbs([(1 << 512) % oracle.R])[0]  # Montgomery form: v R = mont_mul(R^2, v)
        B.call("ff_ax", B.ptr(r2), ctypes.c_void_p(d_s), ctypes.c_void_p(d_s), ctypes.c_uint(n), dev.stream)
        dev.sync()
        s_mont = dev.download(d_s, (n, 8))
        want = msm(dev, d_b, k, L)
        got = msm(dev, d_b, k, L, host_scalars=True)
        assert np.array_equal(oracle.to_affine(got), oracle.to_affine(want))
        t = oracle.from_mont(oracle.fr_dot(k, s_mont))[0]
        assert np.array_equal(oracle.to_affine(oracle.msm_fold(got))[0], oracle.mul_generator(t))
    finally:
        for p in bufs:
            dev.free(p)


@pytest.mark.parametrize("bases_where,scalars_where", [("pageable", "device"), ("pinned", "device"), ("device", "pinned"), ("pinned", "pageable"),
                                                       ("pageable", "pageable")])
def test_msm_host_inputs_chunked(bellman, oracle, bases_where, scalars_where):
    """Host bases and/or host scalars at 2^20: the call copies them in four
    chunks on a copy stream, each chunk's copy overlapping the accumulation of
    the chunk before (src/msm.cu:160-195).  The per-bit results equal the
    device-input call's as points, and the fold equals the closed form; the h2d
    hooks fire once, after the last chunk."""
    import ctypes

    B, dev = bellman
    L = 20
    n = 1 << L
    s0, d = 0xBC000091, 0x0123456789ABCDEF
    bases = oracle.gen_bases_progression(s0, d, n)
    k = oracle.msm_scalars(0xBC000092, n)
    db, dk = dev.upload(bases), dev.upload(k)
    want = msm(dev, db, k, L)
    hooks = CopyHooks()
    releases = []
    try:
        if bases_where == "device":
            b_addr = db
        else:
            b_addr, keep_b, rel = _host_buffer(B, bases, bases_where == "pinned")
            releases.append(rel)
        if scalars_where == "device":
            k_addr = dk
        else:
            k_addr, keep_k, rel = _host_buffer(B, k, scalars_where == "pinned")
            releases.append(rel)
        d_r = dev.malloc(254 * 96)
        cfg = B.msm_configuration()
        cfg.mem_pool, cfg.stream = dev.pool, dev.stream
        cfg.bases, cfg.scalars, cfg.results, cfg.log_scalars_count = b_addr, k_addr, d_r, L
        hooks.attach(cfg)
        B.call("msm_execute_async", cfg)
        dev.sync()
        got = dev.download(d_r, (254, 24))
        dev.free(d_r)
        assert np.array_equal(oracle.to_affine(got), oracle.to_affine(want))
        t = oracle.from_mont(oracle.fr_dot(k, oracle.fr_progression(s0, d, n)))[0]
        assert np.array_equal(oracle.to_affine(oracle.msm_fold(got))[0], oracle.mul_generator(t))
        assert hooks.wait_calls(1) == [CopyHooks.H2D]
        assert hooks.events_complete()
    finally:
        hooks.close()
        for rel in releases:
            rel()
        dev.free(db)
        dev.free(dk)
