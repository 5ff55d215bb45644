"""GPU parity at the workload sizes the benchmark times, through the C ABI.

C3: 2^24 forward NTTs with a 4x coset extension (all four cosets, the
    three-pass coset path with its 4,096-entry coarse twist table), sampled
    outputs against brute-force polynomial evaluation and the inverse-coset
    round trip; the same for bit-reversed inputs at 2^22.
C4: the batched mix (64 MSMs 2^12..2^20 over one shared 2^20 base set, 32 NTTs
    2^10..2^22) through a captured plan and through the batch calls: every MSM
    fold against the closed form [sum k_i s_i]G, every NTT against the oracle
    (full outputs up to 2^20, sampled evaluations above), and everything equal
    to the same jobs issued one at a time through the plain C ABI.
C5: the served mix (50% MSM 2^14..2^22, 35% NTT 2^12..2^22, 15% PN 4 x 2^20)
    through the continuous-batching server with the same checks, PN against
    the oracle.
MSM per-bit results at the wide windows the plain call uses from 2^22 up
(c = 17, 19, 20, 22), projectively against the oracle.

Bases are the progression [s0 + i d]G (made by the oracle), so a fold's
expected value is [sum k_i (s0 + i d)]G, computed by the oracle's Fr dot
product and its generator multiple."""
import ctypes

import numpy as np
import pytest

import bellman_amd as B
from gpu_helpers import msm_batch

pytestmark = pytest.mark.gpu


def brev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2) if bits else 0


def ntt_call(dev, d_in, d_out, L, bitrev=False, inverse=False, log_ext=0, coset=0):
    cfg = B.ntt_configuration()
    cfg.mem_pool, cfg.stream, cfg.inputs, cfg.outputs = dev.pool, dev.stream, d_in, d_out
    cfg.log_values_count, cfg.bit_reversed_inputs, cfg.inverse = L, bitrev, inverse
    cfg.log_extension_degree, cfg.coset_index = log_ext, coset
    B.call("ntt_execute_async", cfg)


def msm_call(dev, d_bases, d_scalars, d_results, L):
    cfg = B.msm_configuration()
    cfg.mem_pool, cfg.stream = dev.pool, dev.stream
    cfg.bases, cfg.scalars, cfg.results, cfg.log_scalars_count = d_bases, d_scalars, d_results, L
    B.call("msm_execute_async", cfg)


def pn_call(dev, d_idx, d_ks, d_out, cols, log_rows):
    cfg = B.generate_permutation_polynomials_configuration(dev.pool, dev.stream, d_idx, d_ks, d_out, cols, log_rows)
    B.call("pn_generate_permutation_polynomials", cfg)


def coset_points(oracle, L, log_ext, coset, ks):
    """g w_{2^(L+e)}^coset w_{2^L}^k for each k (w_{2^L}^k without a coset), Montgomery limbs."""
    R = oracle.R
    w28 = pow(7, (R - 1) >> 28, R)
    w_big = pow(w28, 1 << (28 - L - log_ext), R)
    w = pow(w28, 1 << (28 - L), R)
    rho = 7 * pow(w_big, coset, R) % R if log_ext else 1
    return oracle.to_mont([rho * pow(w, k, R) % R for k in ks])


def check_ntt_sampled(oracle, coeffs, out, L, samples, log_ext=0, coset=0, bitrev_out=True):
    """out (natural -> bit-reversed if bitrev_out, else natural) holds P(rho w^k) at brev(k) (or k)."""
    zs = coset_points(oracle, L, log_ext, coset, samples)
    vals = oracle.poly_eval_many(coeffs, zs)
    for k, v in zip(samples, vals):
        pos = brev(k, L) if bitrev_out else k
        assert np.array_equal(out[pos], v), (L, log_ext, coset, k)


@pytest.mark.timeout(600)
def test_c3_lde_2_24_four_cosets(bellman, oracle):
    """2^24 coefficients, log_extension_degree 2, cosets 0..3 (2^26 outputs in
    all): sampled outputs equal P(g w_{2^26}^c w_{2^24}^k) at position brev(k),
    and the inverse coset transform of each output returns the input."""
    B_, dev = bellman
    L, e = 24, 2
    n = 1 << L
    x = oracle.random_fr(0xBC000003, n)
    d_x = dev.upload(x)
    d_y = dev.malloc(32 * n)
    d_z = dev.malloc(32 * n)
    samples = [0, 1, 4097, 0x9E3779 & (n - 1), n - 1]
    try:
        for c in range(1 << e):
            ntt_call(dev, d_x, d_y, L, log_ext=e, coset=c)
            dev.sync()
            y = dev.download(d_y, (n, 8))
            check_ntt_sampled(oracle, x, y, L, samples, e, c)
            ntt_call(dev, d_y, d_z, L, bitrev=True, inverse=True, log_ext=e, coset=c)
            dev.sync()
            assert np.array_equal(dev.download(d_z, (n, 8)), x), c
    finally:
        for p in (d_x, d_y, d_z):
            dev.free(p)


@pytest.mark.timeout(300)
def test_coset_bit_reversed_inputs_2_22(bellman, oracle):
    """Bit-reversed coefficients in, natural order out, cosets 0 and 3 of a 4x
    extension at 2^22 (three passes, the twist on the first pass of the
    Gentleman-Sande order): output k is P(g w_{2^24}^c w_{2^22}^k) for the
    polynomial whose coefficient i is input brev(i); the inverse returns it."""
    B_, dev = bellman
    L, e = 22, 2
    n = 1 << L
    x = oracle.random_fr(0xBC000033, n)
    coeffs = oracle.bit_reverse(x, L)
    d_x = dev.upload(x)
    d_y = dev.malloc(32 * n)
    d_z = dev.malloc(32 * n)
    try:
        for c in (0, 3):
            ntt_call(dev, d_x, d_y, L, bitrev=True, log_ext=e, coset=c)
            dev.sync()
            y = dev.download(d_y, (n, 8))
            check_ntt_sampled(oracle, coeffs, y, L, [0, 1, 2, 12345, n - 1], e, c, bitrev_out=False)
            ntt_call(dev, d_y, d_z, L, inverse=True, log_ext=e, coset=c)
            dev.sync()
            assert np.array_equal(dev.download(d_z, (n, 8)), x), c
    finally:
        for p in (d_x, d_y, d_z):
            dev.free(p)


@pytest.fixture(scope="module")
def progression_bases(bellman, oracle):
    """2^22 bases [s0 + i d]G on the device, and s_i (Montgomery Fr) on the host."""
    _, dev = bellman
    s0, d = 0xBC000002, 0x1F2E3D4C5B6A7988
    n = 1 << 22
    bases = oracle.gen_bases_progression(s0, d, n)
    db = dev.upload(bases)
    yield db, bases, oracle.fr_progression(s0, d, n)
    dev.free(db)


def expected_fold(oracle, s_mont, k_mont):
    """[sum k_i s_i]G as canonical affine Montgomery limbs."""
    t = oracle.from_mont(oracle.fr_dot(k_mont, s_mont[: k_mont.shape[0]]))[0]
    return oracle.mul_generator(t)


def fold(oracle, res):
    return oracle.to_affine(oracle.msm_fold(res))[0]


@pytest.mark.parametrize("L,c,piece", [(12, 17, 16), (12, 19, 32), (12, 20, 64), (10, 22, 8), (11, 18, 128)])
def test_msm_per_bit_wide_windows(bellman, oracle, progression_bases, L, c, piece):
    """Per-bit results at the window widths of the plain 2^22..2^26 calls,
    through a batch job with the width set: projectively equal to the oracle's
    per-bit results for that width, and folded to the closed form."""
    _, dev = bellman
    db, bases, s_mont = progression_bases
    k = oracle.msm_scalars(0xBC000090 + c, 1 << L)
    ds, dr = dev.upload(k), dev.malloc(254 * 96)
    try:
        msm_batch(dev, [B.msm_job(db, ds, dr, L, c, piece)])
        got = dev.download(dr, (254, 24))
    finally:
        dev.free(ds)
        dev.free(dr)
    exp = oracle.msm(bases[: 1 << L], k, L, c)
    assert np.array_equal(oracle.to_affine(got), oracle.to_affine(exp)), (L, c)
    assert np.array_equal(fold(oracle, got), expected_fold(oracle, s_mont, k))


def _check_ntt_output(oracle, x, L, got):
    """Forward natural -> bit-reversed output: the oracle's in full up to 2^20,
    sampled evaluations above."""
    if L <= 20:
        assert np.array_equal(got, oracle.ntt(x, L)), L
    else:
        check_ntt_sampled(oracle, x, got, L, [0, 1, (1 << L) - 1, 0x5A5A5 & ((1 << L) - 1)])


@pytest.mark.timeout(900)
def test_c4_batched_mix_plan_and_batch_calls(bellman, oracle, progression_bases):
    """64 MSMs with log sizes uniform in 12..20 over the shared bases (prefixes
    of the 2^20 set) and 32 forward NTTs with log sizes uniform in 10..22, the
    mix the benchmark times: a captured plan (replayed twice) and the batch
    entry points give the plain calls' NTT bytes and MSM folds, the folds equal
    the closed form and the NTT outputs the oracle's."""
    _, dev = bellman
    db, bases, s_mont = progression_bases
    rng = np.random.RandomState(0xBC000004 & 0xFFFFFFFF)
    msm_logs = [int(v) for v in rng.randint(12, 21, 64)]
    ntt_logs = [int(v) for v in rng.randint(10, 23, 32)]
    bufs = []

    def alloc(nbytes):
        p = dev.malloc(nbytes)
        bufs.append(p)
        return p

    try:
        ks = [oracle.msm_scalars(0xBC000040 + j, 1 << L) for j, L in enumerate(msm_logs)]
        d_k = []
        for k in ks:
            d_k.append(dev.upload(k))
            bufs.append(d_k[-1])
        r_plain = [alloc(254 * 96) for _ in msm_logs]
        r_plan = [alloc(254 * 96) for _ in msm_logs]
        r_batch = [alloc(254 * 96) for _ in msm_logs]
        xs = [oracle.random_fr(0xBC000070 + j, 1 << L) for j, L in enumerate(ntt_logs)]
        x_in = []
        for x in xs:
            x_in.append(dev.upload(x))
            bufs.append(x_in[-1])
        y_plain = [alloc(32 << L) for L in ntt_logs]
        y_plan = [alloc(32 << L) for L in ntt_logs]
        y_batch = [alloc(32 << L) for L in ntt_logs]

        for j, L in enumerate(msm_logs):
            msm_call(dev, db, d_k[j], r_plain[j], L)
            dev.sync()
        for j, L in enumerate(ntt_logs):
            ntt_call(dev, x_in[j], y_plain[j], L)
            dev.sync()
        print("c4: plain calls done", flush=True)

        def jobs(r, y):
            m = [B.msm_job(db, d_k[j], r[j], L) for j, L in enumerate(msm_logs)]
            t = [B.ntt_job(x_in[j], y[j], L) for j, L in enumerate(ntt_logs)]
            return B.array_of(B.bc_amd_msm_job, m), B.array_of(B.bc_amd_ntt_job, t)

        marr, narr = jobs(r_plan, y_plan)
        plan = B.bc_amd_plan()
        B.call("bc_amd_plan_create", ctypes.byref(plan), marr, ctypes.c_uint(64), narr, ctypes.c_uint(32))
        try:
            for replay in range(2):
                for p in r_plan:
                    B.call("bc_memset", ctypes.c_void_p(p), ctypes.c_int(0), ctypes.c_size_t(254 * 96))
                B.call("bc_amd_plan_launch", plan, dev.stream)
                dev.sync()
                for j, L in enumerate(ntt_logs):
                    assert np.array_equal(dev.download(y_plan[j], (1 << L, 8)), dev.download(y_plain[j], (1 << L, 8))), (replay, L)
                if replay == 0:
                    plan_res = [dev.download(p, (254, 24)) for p in r_plan]
                else:
                    for j, p in enumerate(r_plan):
                        assert np.array_equal(dev.download(p, (254, 24)), plan_res[j]), (replay, j)
        finally:
            B.call("bc_amd_plan_destroy", plan)

        print("c4: plan replays done", flush=True)
        marr, narr = jobs(r_batch, y_batch)
        B.call("bc_amd_msm_batch_async", marr, ctypes.c_uint(64), dev.pool, dev.stream)
        B.call("bc_amd_ntt_batch_async", narr, ctypes.c_uint(32), dev.pool, dev.stream)
        dev.sync()
        for j, L in enumerate(msm_logs):
            plain = dev.download(r_plain[j], (254, 24))
            batch = dev.download(r_batch[j], (254, 24))
            # the plan and the batch call run the same tuned width: identical per-bit results
            assert np.array_equal(batch, plan_res[j]), j
            want = expected_fold(oracle, s_mont, ks[j])
            assert np.array_equal(fold(oracle, plain), want), (j, L)
            assert np.array_equal(fold(oracle, batch), want), (j, L)
        for j, L in enumerate(ntt_logs):
            got = dev.download(y_batch[j], (1 << L, 8))
            assert np.array_equal(got, dev.download(y_plain[j], (1 << L, 8))), (j, L)
            _check_ntt_output(oracle, xs[j], L, got)
    finally:
        for p in bufs:
            dev.free(p)


@pytest.mark.timeout(900)
def test_c5_served_mix(bellman, oracle, progression_bases):
    """120 jobs drawn as the served benchmark draws them (50% MSM 2^14..2^22,
    35% NTT 2^12..2^22, 15% PN 4 x 2^20 with random cycles of length 1..64),
    all submitted to the server at once: every output equals the job's plain
    call (NTT and PN bytes, MSM fold), the MSM folds equal the closed form, the
    NTT outputs the oracle's (full or sampled) and the PN targets the oracle's."""
    from bench import pn_indexes

    _, dev = bellman
    db, bases, s_mont = progression_bases
    rng = np.random.RandomState(0xBC000005 & 0xFFFFFFFF)
    k_all = oracle.msm_scalars(0xBC000006, (1 << 22) + (1 << 21))
    x_all = oracle.random_fr(0xBC000007, 1 << 23)
    pn_idx = [pn_indexes(0xBC000008 + t, 4, 20) for t in range(2)]
    pn_ks = oracle.random_fr(0xBC00000A, 4)
    bufs = []

    def keep(p):
        bufs.append(p)
        return p

    try:
        d_kall, d_xall = keep(dev.upload(k_all)), keep(dev.upload(x_all))
        d_pn_idx = [keep(dev.upload(i)) for i in pn_idx]
        d_pn_ks = keep(dev.upload(pn_ks))
        jobs = []  # (kind, params, plain output, served output)
        for j in range(120):
            u = rng.randint(100)
            if u < 50:
                L, off = int(rng.randint(14, 23)), int(rng.randint(0, 1 << 21))
                jobs.append(("msm", (L, off), keep(dev.malloc(254 * 96)), keep(dev.malloc(254 * 96))))
            elif u < 85:
                L = int(rng.randint(12, 23))
                off = int(rng.randint(0, (1 << 23) - (1 << L) + 1))
                jobs.append(("ntt", (L, off), keep(dev.malloc(32 << L)), keep(dev.malloc(32 << L))))
            else:
                jobs.append(("pn", (j % 2,), keep(dev.malloc(32 * (4 << 20))), keep(dev.malloc(32 * (4 << 20)))))
        for kind, prm, out, _ in jobs:
            if kind == "msm":
                msm_call(dev, db, d_kall + 32 * prm[1], out, prm[0])
            elif kind == "ntt":
                ntt_call(dev, d_xall + 32 * prm[1], out, prm[0])
            else:
                pn_call(dev, d_pn_idx[prm[0]], d_pn_ks, out, 4, 20)
            dev.sync()
        print("c5: plain calls done", flush=True)
        srv = B.Server(hbm_budget_bytes=48 << 30, max_in_flight=3)
        try:
            tickets = []
            for kind, prm, _, out in jobs:
                job = B.bc_amd_job()
                if kind == "msm":
                    job.kind, job.msm = B.BC_AMD_JOB_MSM, B.msm_job(db, d_kall + 32 * prm[1], out, prm[0])
                elif kind == "ntt":
                    job.kind, job.ntt = B.BC_AMD_JOB_NTT, B.ntt_job(d_xall + 32 * prm[1], out, prm[0])
                else:
                    job.kind, job.pn = B.BC_AMD_JOB_PN, B.bc_amd_pn_job(d_pn_idx[prm[0]], d_pn_ks, out, 4, 20)
                tickets.append(srv.submit(job))
            stats = [srv.wait(t) for t in tickets]
        finally:
            srv.close()
        print("c5: served jobs done", flush=True)
        assert all(s.status == 0 for s in stats)
        assert max(s.batch_jobs for s in stats) > 1
        pn_exp = {}
        for kind, prm, plain, served in jobs:
            if kind == "msm":
                L, off = prm
                a, b = dev.download(plain, (254, 24)), dev.download(served, (254, 24))
                want = expected_fold(oracle, s_mont, k_all[off:off + (1 << L)])
                assert np.array_equal(fold(oracle, a), want), prm
                assert np.array_equal(fold(oracle, b), want), prm
            elif kind == "ntt":
                L, off = prm
                a, b = dev.download(plain, (1 << L, 8)), dev.download(served, (1 << L, 8))
                assert np.array_equal(a, b), prm
                _check_ntt_output(oracle, np.ascontiguousarray(x_all[off:off + (1 << L)]), L, b)
            else:
                a, b = dev.download(plain, (4 << 20, 8)), dev.download(served, (4 << 20, 8))
                assert np.array_equal(a, b), prm
                t = prm[0]
                if t not in pn_exp:
                    pn_exp[t] = oracle.pn(pn_idx[t], pn_ks, 4, 20)
                assert np.array_equal(b, pn_exp[t]), prm
    finally:
        for p in bufs:
            dev.free(p)


def test_msm_backs_off_when_the_pool_cannot_hold_one_block(bellman, oracle, progression_bases):
    """Out-of-memory back-off (src/msm.cu:442-459; the pool-filling pattern of
    tests/msm_test.cu:144-153): device memory is filled until a fresh pool can
    no longer supply the temporaries of a 2^22 MSM in one block, but can
    supply those of a half-size chunk.  msm_execute_async re-plans with
    smaller chunks and completes; the fold equals the closed form and the
    per-bit results equal the unconstrained call's as points."""
    _, dev = bellman
    db, bases, s_mont = progression_bases
    L = 22
    k = oracle.msm_scalars(0xBC0000A0, 1 << L)
    dk, dr = dev.upload(k), dev.malloc(254 * 96)
    free_want = None
    pool = B.bc_mem_pool()
    B.call("bc_mem_pool_create", ctypes.byref(pool), ctypes.c_int(0))
    dummies = []
    try:
        msm_call(dev, db, dk, dr, L)
        dev.sync()
        want = dev.download(dr, (254, 24))
        fps = []
        for log_n in (L, L - 1):
            job = B.bc_amd_job()
            job.kind = B.BC_AMD_JOB_MSM
            piece = 32  # the plain call's chunk of sorted entries at 2^21 and 2^22
            job.msm = B.msm_job(db, dk, dr, log_n, oracle.msm_window_bits(L), piece)
            fp = ctypes.c_size_t()
            B.call("bc_amd_job_footprint", ctypes.byref(job), ctypes.byref(fp))
            fps.append(fp.value)
        assert fps[1] < fps[0]
        free_want = (fps[0] + fps[1]) // 2
        # fill device memory (the free figure the runtime reports is not exact
        # on every box): blocks of 1 GiB until an allocation fails, two of them
        # given back and refilled with 64 MiB blocks, then just enough 64 MiB
        # blocks given back for the half-size chunk (< 64 MiB more stays free)
        def fill(size):
            while True:
                blk = ctypes.c_void_p()
                if B.lib().bc_malloc(ctypes.byref(blk), ctypes.c_size_t(size)) != B.bc_success:
                    return
                dummies.append((blk, size))

        fill(1 << 30)
        for _ in range(2):
            if dummies:
                B.call("bc_free", dummies.pop()[0])
        fill(64 << 20)
        ctypes.CDLL("libamdhip64.so").hipGetLastError()
        given = 0
        while given < free_want and dummies and dummies[-1][1] == 64 << 20:
            B.call("bc_free", dummies.pop()[0])
            given += 64 << 20
        assert given >= free_want, "device memory could not be filled in 64 MiB steps"
        # the single block of the unchunked call no longer fits
        probe = ctypes.c_void_p()
        code = B.lib().bc_malloc_from_pool_async(ctypes.byref(probe), ctypes.c_size_t(fps[0]), pool, dev.stream)
        assert code == B.bc_error_memory_allocation, "memory could not be constrained (probe returned %d)" % code
        # the failed probe is the runtime's last error; clear it so that only the call's own allocation fails
        ctypes.CDLL("libamdhip64.so").hipGetLastError()
        B.call("bc_memset", ctypes.c_void_p(dr), ctypes.c_int(0), ctypes.c_size_t(254 * 96))
        cfg = B.msm_configuration()
        cfg.mem_pool, cfg.stream = pool, dev.stream
        cfg.bases, cfg.scalars, cfg.results, cfg.log_scalars_count = db, dk, dr, L
        B.call("msm_execute_async", cfg)
        dev.sync()
        got = dev.download(dr, (254, 24))
    finally:
        dev.sync()
        for blk, _ in dummies:
            B.call("bc_free", blk)
        B.call("bc_mem_pool_destroy", pool)
        dev.free(dk)
        dev.free(dr)
    assert np.array_equal(oracle.to_affine(got), oracle.to_affine(want))
    assert np.array_equal(fold(oracle, got), expected_fold(oracle, s_mont, k))
