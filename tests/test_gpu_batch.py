"""Batched MSM / NTT, hipGraph plans and the serving loop, through the C ABI.

Every batched result is checked two ways: against the CPU oracle (per-bit MSM
results projectively for the job's window width, the MSM fold against the
closed form [sum k_i s_i]G, NTT outputs bit for bit) and against the same job
run alone, byte for byte (a job's launch share does not depend on the batch)."""
import ctypes
import time

import numpy as np
import pytest

import bellman_amd as B
from gpu_helpers import msm, msm_batch, ntt, ntt_batch, pn, pn_cycles

pytestmark = pytest.mark.gpu


def closed_form(oracle, s_ints, k_mont):
    ks = oracle.from_mont(k_mont)
    return oracle.mul_generator(sum(a * b for a, b in zip(ks, s_ints)) % oracle.R)


@pytest.fixture(scope="module")
def shared_bases(bellman, oracle):
    """2^12 bases [s0 + i d]G on the device, with their discrete logs."""
    _, dev = bellman
    s0, d = 0xBC000041, 0x3C6EF372FE94F82B
    n = 1 << 12
    bases = oracle.gen_bases_progression(s0, d, n)
    db = dev.upload(bases)
    yield db, bases, [(s0 + i * d) % oracle.R for i in range(n)]
    dev.free(db)


# (log n, window bits, piece): sizes uneven, widths from 3 to 13, pieces 1..64
MSM_CASES = [(0, 3, 32), (3, 4, 1), (7, 5, 8), (10, 8, 32), (12, 13, 64), (5, 6, 2), (12, 9, 16), (9, 11, 32), (1, 16, 4)]


def test_msm_batch_matches_solo_and_oracle(bellman, oracle, shared_bases):
    _, dev = bellman
    db, bases, logs = shared_bases
    jobs, scal, res = [], [], []
    for j, (L, c, piece) in enumerate(MSM_CASES):
        k = oracle.msm_scalars(0xBC000042 + j, 1 << L)
        scal.append(k)
        ds, dr = dev.upload(k), dev.malloc(254 * 96)
        res.append(dr)
        jobs.append(B.msm_job(db, ds, dr, L, c, piece))
    msm_batch(dev, jobs)
    batched = [dev.download(r, (254, 24)) for r in res]
    for j, (L, c, piece) in enumerate(MSM_CASES):
        msm_batch(dev, [jobs[j]])
        solo = dev.download(res[j], (254, 24))
        assert np.array_equal(batched[j], solo), (L, c, piece)
        exp = oracle.msm(bases[: 1 << L], scal[j], L, c)
        assert np.array_equal(oracle.to_affine(batched[j]), oracle.to_affine(exp)), (L, c, piece)
        assert np.array_equal(oracle.to_affine(oracle.msm_fold(batched[j]))[0], closed_form(oracle, logs[: 1 << L], scal[j]))
    for job in jobs:
        dev.free(job.scalars)
        dev.free(job.results)


def test_msm_batch_tuned_defaults_fold_like_plain_call(bellman, oracle, shared_bases):
    """window_bits = piece = 0 takes the tuning table; the fold equals the plain call's."""
    _, dev = bellman
    db, bases, logs = shared_bases
    L = 12
    k = oracle.msm_scalars(0xBC000043, 1 << L)
    ds, dr = dev.upload(k), dev.malloc(254 * 96)
    msm_batch(dev, [B.msm_job(db, ds, dr, L)])
    got = dev.download(dr, (254, 24))
    plain = msm(dev, db, k, L)
    assert np.array_equal(oracle.to_affine(oracle.msm_fold(got))[0], oracle.to_affine(oracle.msm_fold(plain))[0])
    assert np.array_equal(oracle.to_affine(oracle.msm_fold(got))[0], closed_form(oracle, logs, k))
    dev.free(ds)
    dev.free(dr)


def test_msm_batch_over_64_jobs_and_per_job_bases(bellman, oracle):
    """70 jobs (two launches of <= 64) over two different base sets."""
    _, dev = bellman
    n = 16
    sets = []
    for t in range(2):
        s = [(0x51 + 97 * t + 13 * i) % oracle.R for i in range(n)]
        sets.append((dev.upload(oracle.gen_bases(s)), s))
    jobs, scal = [], []
    for j in range(70):
        db, _ = sets[j % 2]
        k = oracle.msm_scalars(0xBC000044 + j, n)
        scal.append(k)
        jobs.append(B.msm_job(db, dev.upload(k), dev.malloc(254 * 96), 4, 3 + j % 6, 1 + j % 5))
    msm_batch(dev, jobs)
    for j, job in enumerate(jobs):
        got = dev.download(job.results, (254, 24))
        assert np.array_equal(oracle.to_affine(oracle.msm_fold(got))[0], closed_form(oracle, sets[j % 2][1], scal[j])), j
        dev.free(job.scalars)
        dev.free(job.results)
    for db, _ in sets:
        dev.free(db)


def test_msm_batch_rejects_bad_jobs(bellman):
    _, dev = bellman
    bad = B.array_of(B.bc_amd_msm_job, [B.msm_job(None, None, None, 4)])
    assert B.lib().bc_amd_msm_batch_async(bad, ctypes.c_uint(1), dev.pool, dev.stream) == B.bc_error_invalid_value
    p = dev.malloc(1024)
    bad = B.array_of(B.bc_amd_msm_job, [B.msm_job(p, p, p, 4, 25, 0)])
    assert B.lib().bc_amd_msm_batch_async(bad, ctypes.c_uint(1), dev.pool, dev.stream) == B.bc_error_invalid_value
    bad = B.array_of(B.bc_amd_msm_job, [B.msm_job(p, p, p, 4, 8, 1 << 16)])  # chunk above 65535 entries
    assert B.lib().bc_amd_msm_batch_async(bad, ctypes.c_uint(1), dev.pool, dev.stream) == B.bc_error_invalid_value
    dev.free(p)


BAD_NTT_SHAPES = [
    (20, 0xFFFFFFF0, 0),  # log_n + log_ext wraps below 2^32
    (20, 9, 0),  # extended domain above 2^28
    (4, 2, 4),  # coset index not below 2^log_ext
    (0, 0, 0),
    (29, 0, 0),
    (100, 0, 0),
]


@pytest.mark.parametrize("log_n,log_ext,coset", BAD_NTT_SHAPES)
def test_ntt_entry_points_reject_bad_shapes(bellman, log_n, log_ext, coset):
    """ntt_execute_async, the NTT batch, plan creation and server submission
    all answer bc_error_invalid_value for a transform shape outside 2^1..2^28
    (extended domain included), before sizing or allocating anything."""
    _, dev = bellman
    p = dev.malloc(1024)
    try:
        cfg = B.ntt_configuration()
        cfg.mem_pool, cfg.stream, cfg.inputs, cfg.outputs = dev.pool, dev.stream, p, p
        cfg.log_values_count, cfg.log_extension_degree, cfg.coset_index = log_n, log_ext, coset
        assert B.lib().ntt_execute_async(cfg) == B.bc_error_invalid_value
        job = B.ntt_job(p, p, log_n, log_ext=log_ext, coset=coset)
        jobs = B.array_of(B.bc_amd_ntt_job, [job])
        assert B.lib().bc_amd_ntt_batch_async(jobs, ctypes.c_uint(1), dev.pool, dev.stream) == B.bc_error_invalid_value
        plan = B.bc_amd_plan()
        assert B.lib().bc_amd_plan_create(ctypes.byref(plan), None, ctypes.c_uint(0), jobs, ctypes.c_uint(1)) == B.bc_error_invalid_value
        assert not plan.handle
        srv = B.Server(hbm_budget_bytes=1 << 20)
        try:
            sj = B.bc_amd_job()
            sj.kind = B.BC_AMD_JOB_NTT
            sj.ntt = job
            t = ctypes.c_ulonglong()
            assert B.lib().bc_amd_server_submit(srv.handle, ctypes.byref(sj), ctypes.byref(t)) == B.bc_error_invalid_value
        finally:
            srv.close()
        dev.sync()
    finally:
        dev.free(p)


def test_msm_batch_chunk_edges(bellman, oracle, shared_bases):
    """Chunk lengths at the edges of the accumulation's layout, batched together:
    a chunk longer than the whole job (65535 entries), chunks of one entry, an
    all-zero job (no bucket receives a point), an all-one job (window 0's
    digit-1 bucket holds every point: a heavy bucket of 2^11 one-entry sums)
    and r - 1 (every digit negated).  Each job's per-bit results equal its solo
    run byte for byte, the oracle's projectively, and fold to the closed form."""
    _, dev = bellman
    db, bases, logs = shared_bases
    R = oracle.R
    cases = [(10, 8, 65535, None), (9, 5, 1, None), (8, 6, 16, [0] * 256), (11, 7, 1, [1] * 2048), (7, 9, 3, [R - 1] * 128)]
    jobs, scal = [], []
    for j, (L, c, piece, ints) in enumerate(cases):
        k = oracle.to_mont(ints) if ints is not None else oracle.msm_scalars(0xBC0000A0 + j, 1 << L)
        scal.append(k)
        jobs.append(B.msm_job(db, dev.upload(k), dev.malloc(254 * 96), L, c, piece))
    msm_batch(dev, jobs)
    batched = [dev.download(job.results, (254, 24)) for job in jobs]
    for j, (L, c, piece, _) in enumerate(cases):
        msm_batch(dev, [jobs[j]])
        assert np.array_equal(batched[j], dev.download(jobs[j].results, (254, 24))), (L, c, piece)
        exp = oracle.msm(bases[: 1 << L], scal[j], L, c)
        assert np.array_equal(oracle.to_affine(batched[j]), oracle.to_affine(exp)), (L, c, piece)
        assert np.array_equal(oracle.to_affine(oracle.msm_fold(batched[j]))[0], closed_form(oracle, logs[: 1 << L], scal[j]))
    for job in jobs:
        dev.free(job.scalars)
        dev.free(job.results)


# (log n, bit-reversed in, inverse, log extension, coset)
NTT_CASES = [(1, False, False, 0, 0), (5, True, True, 0, 0), (10, False, False, 2, 3), (12, True, False, 0, 0), (13, False, True, 1, 1),
             (16, False, False, 0, 0), (16, True, True, 2, 2), (18, False, True, 0, 0), (20, True, False, 0, 0), (3, False, True, 3, 5)]


def test_ntt_batch_matches_plain_and_oracle(bellman, oracle):
    _, dev = bellman
    jobs, xs = [], []
    for j, (L, br, inv, le, co) in enumerate(NTT_CASES):
        x = oracle.random_fr(0xBC000045 + j, 1 << L)
        xs.append(x)
        d_in = dev.upload(x)
        d_out = d_in if j % 2 else dev.malloc(32 << L)  # in place and out of place
        jobs.append(B.ntt_job(d_in, d_out, L, br, inv, le, co))
    ntt_batch(dev, jobs)
    for j, (L, br, inv, le, co) in enumerate(NTT_CASES):
        got = dev.download(jobs[j].outputs, (1 << L, 8))
        assert np.array_equal(got, oracle.ntt(xs[j], L, br, inv, le, co)), NTT_CASES[j]
        assert np.array_equal(got, ntt(dev, xs[j], L, br, inv, le, co)), NTT_CASES[j]
        dev.free(jobs[j].inputs)
        if jobs[j].outputs != jobs[j].inputs:
            dev.free(jobs[j].outputs)


def test_ntt_stage_split_leaves_outputs_unchanged(bellman, oracle):
    _, dev = bellman
    # two passes (16 = 6..10 + rest) and three passes (21 = 1..10 + rest in two)
    for L, splits in ((16, (6, 8, 10)), (21, (1, 5, 10))):
        x = oracle.random_fr(0xBC000046 + L, 1 << L)
        exp = oracle.ntt(x, L)
        first = ctypes.c_uint()
        B.call("bc_amd_ntt_get_split", ctypes.c_uint(L), ctypes.byref(first))
        try:
            for split in splits:
                B.call("bc_amd_ntt_set_split", ctypes.c_uint(L), ctypes.c_uint(split))
                assert np.array_equal(ntt(dev, x, L), exp), (L, split)
                assert np.array_equal(ntt(dev, exp, L, bitrev=True, inverse=True), x), (L, split)
        finally:
            B.call("bc_amd_ntt_set_split", ctypes.c_uint(L), first)


def test_plan_replays_match_eager_batches(bellman, oracle, shared_bases):
    """A captured plan replayed 3 times, with eager work between replays."""
    _, dev = bellman
    db, bases, logs = shared_bases
    msm_jobs, scal = [], []
    for j, (L, c) in enumerate([(6, 5), (12, 10), (9, 7)]):
        k = oracle.msm_scalars(0xBC000047 + j, 1 << L)
        scal.append(k)
        msm_jobs.append(B.msm_job(db, dev.upload(k), dev.malloc(254 * 96), L, c, 16))
    ntt_jobs, xs = [], []
    for j, (L, br, inv, le, co) in enumerate([(14, False, False, 1, 1), (8, True, True, 0, 0), (17, False, False, 0, 0)]):
        x = oracle.random_fr(0xBC000048 + j, 1 << L)
        xs.append((x, L, br, inv, le, co))
        ntt_jobs.append(B.ntt_job(dev.upload(x), dev.malloc(32 << L), L, br, inv, le, co))
    msm_batch(dev, msm_jobs)
    eager = [dev.download(j.results, (254, 24)) for j in msm_jobs]
    plan = B.bc_amd_plan()
    B.call("bc_amd_plan_create", ctypes.byref(plan), B.array_of(B.bc_amd_msm_job, msm_jobs), ctypes.c_uint(3),
           B.array_of(B.bc_amd_ntt_job, ntt_jobs), ctypes.c_uint(3))
    ws = ctypes.c_size_t()
    B.call("bc_amd_plan_workspace_bytes", plan, ctypes.byref(ws))
    assert ws.value > 0
    try:
        for replay in range(3):
            for j in msm_jobs:
                B.call("bc_memset", ctypes.c_void_p(j.results), ctypes.c_int(0), ctypes.c_size_t(254 * 96))
            for j in ntt_jobs:
                B.call("bc_memset", ctypes.c_void_p(j.outputs), ctypes.c_int(0), ctypes.c_size_t(32 << j.log_values_count))
            B.call("bc_amd_plan_launch", plan, dev.stream)
            dev.sync()
            for j, job in enumerate(msm_jobs):
                assert np.array_equal(dev.download(job.results, (254, 24)), eager[j]), (replay, j)
            for j, job in enumerate(ntt_jobs):
                x, L, br, inv, le, co = xs[j]
                assert np.array_equal(dev.download(job.outputs, (1 << L, 8)), oracle.ntt(x, L, br, inv, le, co)), (replay, j)
            # eager work on the same stream between replays
            ntt(dev, xs[0][0], 10)
    finally:
        B.call("bc_amd_plan_destroy", plan)
    for j, job in enumerate(msm_jobs):
        assert np.array_equal(oracle.to_affine(oracle.msm_fold(eager[j]))[0], closed_form(oracle, logs[: 1 << job.log_scalars_count], scal[j]))
        dev.free(job.scalars)
        dev.free(job.results)
    for job in ntt_jobs:
        dev.free(job.inputs)
        dev.free(job.outputs)


def _server_jobs(dev, oracle, bases_ptr):
    """A small mixed job list: (bc_amd_job, kind, expected-output getter)."""
    jobs = []
    for j, L in enumerate([8, 12, 5, 10]):
        k = oracle.msm_scalars(0xBC000049 + j, 1 << L)
        job = B.bc_amd_job()
        job.kind = B.BC_AMD_JOB_MSM
        job.msm = B.msm_job(bases_ptr, dev.upload(k), dev.malloc(254 * 96), L)
        jobs.append((job, ("msm", k, L)))
    for j, (L, br, inv, le, co) in enumerate([(12, False, False, 0, 0), (15, True, True, 0, 0), (11, False, False, 2, 1)]):
        x = oracle.random_fr(0xBC00004A + j, 1 << L)
        job = B.bc_amd_job()
        job.kind = B.BC_AMD_JOB_NTT
        job.ntt = B.ntt_job(dev.upload(x), dev.malloc(32 << L), L, br, inv, le, co)
        jobs.append((job, ("ntt", x, (L, br, inv, le, co))))
    for j, (cols, lr) in enumerate([(4, 10), (3, 8)]):
        idx = pn_cycles(cols, lr, 0xBC00004B + j)
        ks = oracle.random_fr(0xBC00004C + j, cols)
        job = B.bc_amd_job()
        job.kind = B.BC_AMD_JOB_PN
        job.pn = B.bc_amd_pn_job(dev.upload(idx), dev.upload(ks), dev.malloc((cols << lr) * 32), cols, lr)
        jobs.append((job, ("pn", idx, ks, cols, lr)))
    return jobs


def _output(dev, job):
    if job.kind == B.BC_AMD_JOB_MSM:
        return dev.download(job.msm.results, (254, 24))
    if job.kind == B.BC_AMD_JOB_NTT:
        return dev.download(job.ntt.outputs, (1 << job.ntt.log_values_count, 8))
    return dev.download(job.pn.target, (job.pn.columns_count << job.pn.log_rows_count, 8))


def _clear(job):
    if job.kind == B.BC_AMD_JOB_MSM:
        p, n = job.msm.results, 254 * 96
    elif job.kind == B.BC_AMD_JOB_NTT:
        p, n = job.ntt.outputs, 32 << job.ntt.log_values_count
    else:
        p, n = job.pn.target, 32 * (job.pn.columns_count << job.pn.log_rows_count)
    B.call("bc_memset", ctypes.c_void_p(p), ctypes.c_int(0), ctypes.c_size_t(n))


def test_server_results_same_alone_and_in_mix(bellman, oracle, shared_bases):
    _, dev = bellman
    db, bases, logs = shared_bases
    jobs = _server_jobs(dev, oracle, db)
    srv = B.Server(hbm_budget_bytes=1 << 30, max_in_flight=2)
    try:
        alone = []
        for job, _ in jobs:
            _clear(job)
            st = srv.wait(srv.submit(job))
            assert st.status == 0 and st.batch_jobs == 1
            alone.append(_output(dev, job))
        for job, _ in jobs:
            _clear(job)
        tickets = [srv.submit(job) for job, _ in jobs]
        stats = [srv.wait(t) for t in tickets]
        assert all(s.status == 0 for s in stats)
        assert max(s.batch_jobs for s in stats) > 1  # jobs really shared batches
        for s in stats:
            assert s.submit_ns <= s.start_ns <= s.enqueued_ns <= s.done_ns
            assert s.slot < 2
        for (job, exp), a in zip(jobs, alone):
            got = _output(dev, job)
            assert np.array_equal(got, a), exp[0]
            if exp[0] == "msm":
                _, k, L = exp
                assert np.array_equal(oracle.to_affine(oracle.msm_fold(got))[0], closed_form(oracle, logs[: 1 << L], k))
            elif exp[0] == "ntt":
                _, x, (L, br, inv, le, co) = exp
                assert np.array_equal(got, oracle.ntt(x, L, br, inv, le, co))
            else:
                _, idx, ks, cols, lr = exp
                assert np.array_equal(got, oracle.pn(idx, ks, cols, lr))
                assert np.array_equal(got, pn(dev, idx, ks, cols, lr))
    finally:
        srv.close()


def test_footprint_bounds_exact_workspace(bellman, oracle, shared_bases):
    """bc_amd_job_footprint (the admission estimate, and the block a plain call
    takes from its pool) is at least the exact workspace of the job run alone;
    for MSM and NTT jobs that workspace is also what a single-job plan holds."""
    _, dev = bellman
    db, _, _ = shared_bases
    k = dev.upload(oracle.msm_scalars(0xBC000050, 1 << 12))
    r = dev.malloc(254 * 96)
    x = dev.upload(oracle.random_fr(0xBC000051, 1 << 12))
    y = dev.malloc(32 << 12)
    idx = dev.upload(pn_cycles(4, 10, 0xBC000052))
    ks = dev.upload(oracle.random_fr(0xBC000053, 4))
    t = dev.malloc(32 * (4 << 10))
    cases = [("msm", B.msm_job(db, k, r, L, c, piece)) for L, c, piece in [(12, 13, 64), (12, 4, 1), (10, 16, 8), (0, 3, 32), (12, 9, 128)]]
    cases += [("ntt", B.ntt_job(x, y, L, br, inv, le, co)) for L, br, inv, le, co in [(12, False, False, 0, 0), (10, True, True, 2, 3), (12, False, True, 0, 0)]]
    cases += [("pn", B.bc_amd_pn_job(idx, ks, t, cols, lr)) for cols, lr in [(4, 10), (3, 10), (4, 6)]]
    try:
        for kind, j in cases:
            job = B.bc_amd_job()
            exact, fp = ctypes.c_size_t(), ctypes.c_size_t()
            if kind == "msm":
                job.kind, job.msm = B.BC_AMD_JOB_MSM, j
            elif kind == "ntt":
                job.kind, job.ntt = B.BC_AMD_JOB_NTT, j
            else:
                job.kind, job.pn = B.BC_AMD_JOB_PN, j
            B.call("bc_amd_job_workspace_bytes", ctypes.byref(job), ctypes.byref(exact))
            B.call("bc_amd_job_footprint", ctypes.byref(job), ctypes.byref(fp))
            assert fp.value >= exact.value, (kind, fp.value, exact.value)
            if kind == "pn":
                continue
            plan, ws = B.bc_amd_plan(), ctypes.c_size_t()
            if kind == "msm":
                B.call("bc_amd_plan_create", ctypes.byref(plan), B.array_of(B.bc_amd_msm_job, [j]), ctypes.c_uint(1), None, ctypes.c_uint(0))
            else:
                B.call("bc_amd_plan_create", ctypes.byref(plan), None, ctypes.c_uint(0), B.array_of(B.bc_amd_ntt_job, [j]), ctypes.c_uint(1))
            try:
                B.call("bc_amd_plan_workspace_bytes", plan, ctypes.byref(ws))
            finally:
                B.call("bc_amd_plan_destroy", plan)
            assert ws.value == exact.value, (kind, ws.value, exact.value)
    finally:
        for p in (k, r, x, y, idx, ks, t):
            dev.free(p)


def test_server_tight_budget_runs_everything(bellman, oracle, shared_bases):
    """A budget smaller than any job: jobs run one batch at a time, all complete."""
    _, dev = bellman
    db, _, logs = shared_bases
    jobs = _server_jobs(dev, oracle, db)
    srv = B.Server(hbm_budget_bytes=4096, max_in_flight=3)
    try:
        tickets = [srv.submit(job) for job, _ in jobs]
        stats = [srv.wait(t) for t in tickets]
        assert all(s.status == 0 and s.batch_jobs == 1 for s in stats)
        for (job, exp) in jobs:
            if exp[0] == "ntt":
                _, x, (L, br, inv, le, co) = exp
                assert np.array_equal(_output(dev, job), oracle.ntt(x, L, br, inv, le, co))
    finally:
        srv.close()


def test_server_shares_workspace_between_batches(bellman, oracle, shared_bases):
    """A budget of about two jobs' footprints and three slots: batches take
    workspace regions side by side, wait for one another when the workspace
    is full, and every output equals the job's output when run alone."""
    _, dev = bellman
    db, _, logs = shared_bases
    jobs = _server_jobs(dev, oracle, db) + _server_jobs(dev, oracle, db)
    fps = []
    for job, _ in jobs:
        fp = ctypes.c_size_t()
        B.call("bc_amd_job_footprint", ctypes.byref(job), ctypes.byref(fp))
        fps.append(fp.value)
    srv = B.Server(hbm_budget_bytes=2 * max(fps), max_in_flight=3)
    try:
        alone = []
        for job, _ in jobs:
            _clear(job)
            assert srv.wait(srv.submit(job)).status == 0
            alone.append(_output(dev, job))
        for rnd in range(3):
            for job, _ in jobs:
                _clear(job)
            tickets = [srv.submit(job) for job, _ in jobs]
            stats = [srv.wait(t) for t in tickets]
            assert all(s.status == 0 for s in stats), rnd
            for s in stats:
                assert s.submit_ns <= s.start_ns <= s.enqueued_ns <= s.done_ns and s.slot < 3
            for (job, exp), a in zip(jobs, alone):
                assert np.array_equal(_output(dev, job), a), (rnd, exp[0])
    finally:
        srv.close()


def test_server_rejects_malformed_jobs(bellman):
    srv = B.Server(hbm_budget_bytes=1 << 20)
    try:
        job = B.bc_amd_job()
        job.kind = B.BC_AMD_JOB_PN
        job.pn = B.bc_amd_pn_job(None, None, None, 5, 4)
        t = ctypes.c_ulonglong()
        assert B.lib().bc_amd_server_submit(srv.handle, ctypes.byref(job), ctypes.byref(t)) == B.bc_error_invalid_value
        job.kind = 7
        assert B.lib().bc_amd_server_submit(srv.handle, ctypes.byref(job), ctypes.byref(t)) == B.bc_error_invalid_value
    finally:
        srv.close()


def test_autotune_sets_valid_tuning_and_keeps_results(bellman, oracle, shared_bases, tmp_path):
    _, dev = bellman
    db, bases, logs = shared_bases
    k = oracle.msm_scalars(0xBC00004D, 1 << 12)
    dk = dev.upload(k)
    saved = str(tmp_path / "before.txt")
    B.call("bc_amd_tuning_save", saved.encode())
    try:
        B.call("bc_amd_msm_autotune", ctypes.c_void_p(db), ctypes.c_void_p(dk), ctypes.c_uint(8), ctypes.c_uint(10), dev.stream)
        scratch = dev.malloc(32 << 14)
        B.call("bc_amd_ntt_autotune", ctypes.c_void_p(scratch), ctypes.c_uint(13), ctypes.c_uint(14), dev.stream)
        dev.free(scratch)
        for L in (8, 9, 10):
            t = B.bc_amd_msm_tuning()
            B.call("bc_amd_msm_get_tuning", ctypes.c_uint(L), ctypes.byref(t))
            assert 3 <= t.window_bits <= 22 and t.piece >= 1
            dr = dev.malloc(254 * 96)
            msm_batch(dev, [B.msm_job(db, dk, dr, L)])
            got = dev.download(dr, (254, 24))
            dev.free(dr)
            assert np.array_equal(oracle.to_affine(oracle.msm_fold(got))[0], closed_form(oracle, logs[: 1 << L], k[: 1 << L]))
        for L in (13, 14):
            f = ctypes.c_uint()
            B.call("bc_amd_ntt_get_split", ctypes.c_uint(L), ctypes.byref(f))
            assert 1 <= f.value <= 10 and L - f.value <= 10  # two passes at these sizes
            x = oracle.random_fr(0xBC00004E + L, 1 << L)
            assert np.array_equal(ntt(dev, x, L), oracle.ntt(x, L))
    finally:
        B.call("bc_amd_tuning_load", saved.encode())
        dev.free(dk)


def test_msm_fold_on_device(bellman, oracle, shared_bases):
    """bc_amd_msm_fold equals the host fold + affine normalisation, infinity as (0, 0)."""
    _, dev = bellman
    db, bases, logs = shared_bases
    ptrs, exp = [], []
    for j, L in enumerate([0, 4, 9, 12, 3]):
        k = oracle.msm_scalars(0xBC00004F + j, 1 << L) if j != 4 else oracle.to_mont([0] * 8)
        dk, dr = dev.upload(k), dev.malloc(254 * 96)
        msm_batch(dev, [B.msm_job(db, dk, dr, L)])
        res = dev.download(dr, (254, 24))
        folded = oracle.to_affine(oracle.msm_fold(res))[0]
        assert np.array_equal(folded, closed_form(oracle, logs[: 1 << L], k))
        exp.append(folded if j != 4 else np.zeros(16, dtype=np.uint32))
        ptrs.append(dr)
        dev.free(dk)
    d_p = dev.malloc(64 * len(ptrs))
    B.call("bc_amd_msm_fold", (ctypes.c_void_p * len(ptrs))(*ptrs), ctypes.c_void_p(d_p), ctypes.c_uint(len(ptrs)), dev.stream)
    dev.sync()
    got = dev.download(d_p, (len(ptrs), 16))
    for j in range(len(ptrs)):
        assert np.array_equal(got[j], exp[j]), j
    dev.free(d_p)
    for p in ptrs:
        dev.free(p)
