"""Fixed-base MSM jobs through the C ABI: bc_amd_msm_fixed_base_table and
batch jobs over its tables.

The table is checked entry by entry against [2^(c w) s_i]G computed by the
oracle from the bases' discrete logs.  An MSM over a table puts every window's
signed digits (the reference's digit rule, src/msm_kernels.cu:36-75, restated
by the oracle) into one window of buckets, so its per-bit result j < c is
[sum of the signed 2^(c w) s_i of every digit with bit j set]G and the ones
above are infinity; these are computed with Python integers from the oracle's
digits.  Folds are checked against the closed form [sum k_i s_i]G and against
the plain batch job's fold, and every job is byte-identical alone and in a
batch, a plan or the server."""
import ctypes

import numpy as np
import pytest

import bellman_amd as B
from gpu_helpers import msm_batch

pytestmark = pytest.mark.gpu


def windows(c):
    return (254 - 1) // c + 1


def make_table(dev, d_bases, log_count, c):
    d_t = dev.malloc(windows(c) * (64 << log_count))
    B.call("bc_amd_msm_fixed_base_table", ctypes.c_void_p(d_bases), ctypes.c_uint(log_count), ctypes.c_uint(c), ctypes.c_void_p(d_t), dev.stream)
    dev.sync()
    return d_t


def closed_form(oracle, s_ints, k_mont):
    ks = oracle.from_mont(k_mont)
    return oracle.mul_generator(sum(a * b for a, b in zip(ks, s_ints)) % oracle.R)


def expected_per_bit(oracle, s_ints, k_mont, c):
    """Discrete logs of the 254 per-bit results of a single-window job."""
    R = oracle.R
    bucket = {}
    for s, k in zip(s_ints, k_mont):
        d, sg = oracle.msm_digits(k, c)
        for w in range(len(d)):
            if d[w]:
                v = (s << (c * w)) % R
                bucket[int(d[w])] = (bucket.get(int(d[w]), 0) + (R - v if sg[w] else v)) % R
    logs = [0] * 254
    for dig, v in bucket.items():
        for j in range(c):
            if (dig >> j) & 1:
                logs[j] = (logs[j] + v) % R
    return logs


@pytest.fixture(scope="module")
def based(bellman, oracle):
    """2^10 bases [s_i]G (random s_i) on the device, with their discrete logs."""
    _, dev = bellman
    L = 10
    s = oracle.from_limbs(oracle.random_fr(0xBC0000F1, 1 << L, montgomery=False))
    bases = oracle.gen_bases(s)
    db = dev.upload(bases)
    yield db, bases, s, L
    dev.free(db)


@pytest.mark.parametrize("c", [3, 7, 20])
def test_fixed_base_table_entries(bellman, oracle, based, c):
    _, dev = bellman
    db, bases, s, L = based
    k = 6  # a table over the first 2^6 bases
    d_t = make_table(dev, db, k, c)
    got = dev.download(d_t, (windows(c) << k, 16))
    dev.free(d_t)
    for w in range(windows(c)):
        for i in range(1 << k):
            exp = oracle.mul_generator((s[i] << (c * w)) % oracle.R)
            assert np.array_equal(got[(w << k) + i], exp), (c, w, i)


@pytest.mark.parametrize("c", [5, 13, 16])
def test_fixed_base_msm_per_bit_and_fold(bellman, oracle, based, c):
    """Jobs of 2^0, 2^5 and 2^10 points over one table, batched with a plain job."""
    _, dev = bellman
    db, bases, s, L = based
    d_t = make_table(dev, db, L, c)
    cases = [(0, 2), (5, 1), (10, 16)]
    jobs, scal, res = [], [], []
    for j, (log_n, piece) in enumerate(cases):
        kk = oracle.msm_scalars(0xBC0000F2 + 16 * c + j, 1 << log_n)
        scal.append(kk)
        ds, dr = dev.upload(kk), dev.malloc(254 * 96)
        res.append(dr)
        jobs.append(B.msm_job(d_t, ds, dr, log_n, c, piece, L))
    # a plain job over the bases in the same launch
    kp = oracle.msm_scalars(0xBC0000F9, 1 << 7)
    dsp, drp = dev.upload(kp), dev.malloc(254 * 96)
    jobs.append(B.msm_job(db, dsp, drp, 7, 9, 8))
    try:
        msm_batch(dev, jobs)
        batched = [dev.download(r, (254, 24)) for r in res]
        plain = dev.download(drp, (254, 24))
        assert np.array_equal(oracle.to_affine(plain), oracle.to_affine(oracle.msm(bases[: 1 << 7], kp, 7, 9)))
        for j, (log_n, piece) in enumerate(cases):
            n = 1 << log_n
            got = oracle.to_affine(batched[j])
            logs = expected_per_bit(oracle, s[:n], scal[j], c)
            for i in range(254):
                assert np.array_equal(got[i], oracle.mul_generator(logs[i])), (c, log_n, i)
            assert np.array_equal(oracle.to_affine(oracle.msm_fold(batched[j]))[0], closed_form(oracle, s[:n], scal[j]))
            msm_batch(dev, [jobs[j]])
            assert np.array_equal(dev.download(res[j], (254, 24)), batched[j]), (c, log_n)
    finally:
        for job in jobs:
            dev.free(job.scalars)
            dev.free(job.results)
        dev.free(d_t)


def test_fixed_base_jobs_in_plan_and_server(bellman, oracle, based):
    """A plan and the server over fixed-base jobs give the eager batch's bytes."""
    _, dev = bellman
    db, bases, s, L = based
    c = 12
    d_t = make_table(dev, db, L, c)
    logs = [10, 4, 9, 7]
    jobs, res, scal = [], [], []
    for j, log_n in enumerate(logs):
        kk = oracle.msm_scalars(0xBC0000FA + j, 1 << log_n)
        scal.append(kk)
        ds, dr = dev.upload(kk), dev.malloc(254 * 96)
        res.append(dr)
        jobs.append(B.msm_job(d_t, ds, dr, log_n, c, 0, L))
    try:
        msm_batch(dev, jobs)
        eager = [dev.download(r, (254, 24)) for r in res]
        for r in res:
            B.call("bc_memset", ctypes.c_void_p(r), ctypes.c_int(0), ctypes.c_size_t(254 * 96))
        plan = B.bc_amd_plan()
        marr = B.array_of(B.bc_amd_msm_job, jobs)
        B.call("bc_amd_plan_create", ctypes.byref(plan), marr, ctypes.c_uint(len(jobs)), None, ctypes.c_uint(0))
        for _ in range(3):
            B.call("bc_amd_plan_launch", plan, dev.stream)
            dev.sync()
            assert all(np.array_equal(dev.download(r, (254, 24)), e) for r, e in zip(res, eager))
        B.call("bc_amd_plan_destroy", plan)
        for r in res:
            B.call("bc_memset", ctypes.c_void_p(r), ctypes.c_int(0), ctypes.c_size_t(254 * 96))
        srv = B.bc_amd_server()
        B.call("bc_amd_server_create", ctypes.byref(srv), ctypes.c_size_t(4 << 30), ctypes.c_uint(2))
        tickets = []
        for job in jobs:
            bj = B.bc_amd_job()
            bj.kind = B.BC_AMD_JOB_MSM
            bj.msm = job
            t = ctypes.c_ulonglong()
            B.call("bc_amd_server_submit", srv, ctypes.byref(bj), ctypes.byref(t))
            tickets.append(t)
        for t in tickets:
            st = B.bc_amd_job_stats()
            B.call("bc_amd_server_wait", srv, t, ctypes.byref(st))
            assert st.status == 0
        B.call("bc_amd_server_destroy", srv)
        assert all(np.array_equal(dev.download(r, (254, 24)), e) for r, e in zip(res, eager))
        for j, log_n in enumerate(logs):
            assert np.array_equal(oracle.to_affine(oracle.msm_fold(eager[j]))[0], closed_form(oracle, s[: 1 << log_n], scal[j]))
    finally:
        for job in jobs:
            dev.free(job.scalars)
            dev.free(job.results)
        dev.free(d_t)


def test_fixed_base_rejections(bellman, oracle, based):
    _, dev = bellman
    db, bases, s, L = based
    d_t = make_table(dev, db, 4, 8)
    kk = oracle.msm_scalars(0xBC0000FF, 1 << 5)
    ds, dr = dev.upload(kk), dev.malloc(254 * 96)
    try:
        lib = B.lib()
        for job in (B.msm_job(d_t, ds, dr, 5, 8, 0, 4),   # more points than the table holds
                    B.msm_job(d_t, ds, dr, 4, 0, 0, 4)):  # the table's width not named
            arr = B.array_of(B.bc_amd_msm_job, [job])
            assert lib.bc_amd_msm_batch_async(arr, ctypes.c_uint(1), dev.pool, dev.stream) == B.bc_error_invalid_value
        for c, k in ((2, 4), (25, 4), (3, 25), (8, 29)):
            assert lib.bc_amd_msm_fixed_base_table(ctypes.c_void_p(db), ctypes.c_uint(k), ctypes.c_uint(c), ctypes.c_void_p(d_t),
                                                   dev.stream) == B.bc_error_invalid_value, (c, k)
        dev.sync()
    finally:
        for p in (ds, dr, d_t):
            dev.free(p)


@pytest.mark.timeout(300)
def test_fixed_base_c4_sizes_against_plain_folds(bellman, oracle):
    """The batched mix's sizes: 2^20 bases, one table (c = 16), jobs of 2^12..2^20
    points; each fold equals the plain call's fold and the closed form."""
    _, dev = bellman
    L, c = 20, 16
    s0, d = 0xBC0000E1, 0x9E3779B97F4A7C15
    bases = oracle.gen_bases_progression(s0, d, 1 << L)
    db = dev.upload(bases)
    d_t = make_table(dev, db, L, c)
    logs = [20, 12, 16, 13, 18]
    jobs, res, plain, scal = [], [], [], []
    try:
        for j, log_n in enumerate(logs):
            kk = oracle.msm_scalars(0xBC0000E2 + j, 1 << log_n)
            scal.append(kk)
            ds, dr, dp = dev.upload(kk), dev.malloc(254 * 96), dev.malloc(254 * 96)
            res.append(dr)
            plain.append(dp)
            jobs.append(B.msm_job(d_t, ds, dr, log_n, c, 0, L))
        msm_batch(dev, jobs)
        msm_batch(dev, [B.msm_job(db, job.scalars, dp, job.log_scalars_count) for job, dp in zip(jobs, plain)])
        arr = (ctypes.c_void_p * (2 * len(jobs)))(*(res + plain))
        d_f = dev.malloc(64 * 2 * len(jobs))
        B.call("bc_amd_msm_fold", arr, ctypes.c_void_p(d_f), ctypes.c_uint(2 * len(jobs)), dev.stream)
        dev.sync()
        folds = dev.download(d_f, (2 * len(jobs), 16))
        dev.free(d_f)
        for j, log_n in enumerate(logs):
            n = 1 << log_n
            ks = oracle.from_mont(scal[j])
            exp = oracle.mul_generator(sum(k * ((s0 + i * d) % oracle.R) for i, k in enumerate(ks)) % oracle.R)
            assert np.array_equal(folds[j], exp), log_n
            assert np.array_equal(folds[len(jobs) + j], exp), log_n
    finally:
        for job in jobs:
            dev.free(job.scalars)
        for p in res + plain + [d_t, db]:
            dev.free(p)
