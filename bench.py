"""Benchmark of the prover hot path on one MI355X.

Headline step: one BN254 G1 MSM of 2^24 points plus one Fr NTT of 2^24
elements (forward, natural -> bit-reversed), both through the bellman-cuda C
ABI with inputs resident in HBM.  Beside it:
  batched: 64 MSMs (2^12..2^20, shared 2^20 bases) + 32 NTTs (2^10..2^22) as one
           captured plan, against the same 96 jobs issued one at a time through
           the plain C ABI with a stream synchronisation after each;
  served:  a mixed stream of MSM (2^14..2^22), NTT (2^12..2^22) and PN
           (4 x 2^20) jobs through the continuous-batching server: capacity,
           p50/p99 latency under Poisson arrivals, and the one-at-a-time FIFO
           under the same arrivals.
The batched and served MSMs run over fixed-base tables of the resident bases
(built once, outside the timed regions; the same forms with the MSMs over the
bases themselves are reported beside them).  Every batched and served result
is checked against its plain-call result (NTT and PN bytes, MSM folded point)
and every MSM fold against the closed form [sum k_i s_i]G, on the device.
Inputs are synthetic and seeded (splitmix64).
Prints one JSON line; progress goes to stderr.
"""
import argparse
import ctypes
import hashlib
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "era-bellman-cuda_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

import bellman_amd as B  # noqa: E402

HBM_PEAK_GBS = 8000.0
# VALU issue peak of MI355X: 256 CUs x 4 SIMD16 x 2.4 GHz = 39.3 T lane-operations
# per second (its 78.6 TFLOP/s FP64 vector rate / 2); a register-only probe of
# v_mad_u64_u32 reached 31 T/s, shown beside it.
VALU_ISSUE_PEAK_T = 39.3
VALU_MAD_PROBE_T = 31.0
R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
MSM_METRIC_LOG = 24


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ---------------------------------------------------------------- synthetic inputs
M64 = (1 << 64) - 1
R_LIMBS = [(R >> (64 * i)) & M64 for i in range(4)]


def splitmix(seed, idx):
    """splitmix64 output number idx of the stream `seed` (vectorised over idx)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (idx.astype(np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform_fr(seed, n):
    """n uniform canonical Fr values as (n, 8) u32 limbs: 254-bit draws, rejected if >= r."""
    out = np.zeros((n, 4), dtype=np.uint64)
    todo = np.arange(n)
    rnd = 0
    while todo.size:
        base = (np.uint64(rnd) << np.uint64(40)) + todo.astype(np.uint64) * np.uint64(4)
        cand = np.stack([splitmix(seed, base + np.uint64(j)) for j in range(4)], axis=1)
        cand[:, 3] &= np.uint64((1 << 62) - 1)
        lt = np.zeros(todo.size, dtype=bool)
        eq = np.ones(todo.size, dtype=bool)
        for j in (3, 2, 1, 0):
            lt |= eq & (cand[:, j] < np.uint64(R_LIMBS[j]))
            eq &= cand[:, j] == np.uint64(R_LIMBS[j])
        out[todo[lt]] = cand[lt]
        todo = todo[~lt]
        rnd += 1
    return out.view(np.uint32).reshape(n, 8)


def msm_scalars_canonical(seed, n):
    """70% uniform in [0, r), 20% below 2^64, 10% in {0, 1} (canonical limbs)."""
    idx = np.arange(n)
    u = splitmix(seed ^ 0x5A5A5A5A, idx) % np.uint64(100)
    small = splitmix(seed ^ 0xA5A5A5A5, idx)
    out = uniform_fr(seed, n)
    out64 = out.view(np.uint64).reshape(n, 4)
    mid = (u >= 70) & (u < 90)
    out64[mid] = 0
    out64[mid, 0] = small[mid]
    tail = u >= 90
    out64[tail] = 0
    out64[tail, 0] = small[tail] & np.uint64(1)
    return out


class Lab:
    """Device, stream, timer and the device-side helpers the checks use."""

    def __init__(self):
        B.call("ff_set_up", ctypes.c_uint(25), ctypes.c_uint(14))
        B.call("ntt_set_up")
        B.call("pn_set_up")
        B.call("msm_set_up")
        self.dev = B.Device(0)
        self.start, self.stop = B.bc_event(), B.bc_event()
        B.call("bc_event_create", ctypes.byref(self.start), ctypes.c_bool(False), ctypes.c_bool(False))
        B.call("bc_event_create", ctypes.byref(self.stop), ctypes.c_bool(False), ctypes.c_bool(False))
        r2 = (1 << 512) % R
        self.r2 = np.array([(r2 >> (32 * i)) & 0xFFFFFFFF for i in range(8)], dtype=np.uint32)
        self.raw_one = np.array([1, 0, 0, 0, 0, 0, 0, 0], dtype=np.uint32)
        mont_one = (1 << 256) % R
        self.mont_one = np.array([(mont_one >> (32 * i)) & 0xFFFFFFFF for i in range(8)], dtype=np.uint32)

    def time_ms(self, fn, reps=1):
        """Events on the library's stream around `reps` calls of fn; ms per call."""
        B.call("bc_event_record", self.start, self.dev.stream)
        for _ in range(reps):
            fn()
        B.call("bc_event_record", self.stop, self.dev.stream)
        B.call("bc_event_synchronize", self.stop)
        ms = ctypes.c_float()
        B.call("bc_event_elapsed_time", ctypes.byref(ms), self.start, self.stop)
        return ms.value / reps

    def to_mont(self, d_canon, d_out, n):
        """Montgomery form on the device: mont_mul(R^2, v) = v R."""
        B.call("ff_ax", B.ptr(self.r2), ctypes.c_void_p(d_canon), ctypes.c_void_p(d_out), ctypes.c_uint(n), self.dev.stream)

    def upload_mont(self, canon):
        d = self.dev.upload(canon)
        self.to_mont(d, d, canon.shape[0])
        return d

    def bases(self, seed, n):
        """[s_i]G for uniform s_i (device), plus the device copy of s_i (canonical)."""
        s = uniform_fr(seed, n)
        d_s = self.dev.upload(s)
        d_b = self.dev.malloc(n * 64)
        B.call("bc_amd_generator_multiples", ctypes.c_void_p(d_s), ctypes.c_void_p(d_b), ctypes.c_uint(n), self.dev.stream)
        return d_b, d_s

    def expected_points(self, d_s, k_ptrs, ns):
        """[sum_i k_i s_i]G for each job (s canonical prefix, k Montgomery), as affine points."""
        dev = self.dev
        m = len(k_ptrs)
        d_t = dev.malloc(32 * m)
        nmax = max(ns)
        d_tmp = dev.malloc(32 * nmax)
        for j, (dk, n) in enumerate(zip(k_ptrs, ns)):
            self.to_mont(d_s, d_tmp, n)
            B.call("ff_x_mul_y", ctypes.c_void_p(d_tmp), ctypes.c_void_p(dk), ctypes.c_void_p(d_tmp), ctypes.c_uint(n), dev.stream)
            cfg = B.ff_poly_evaluate_configuration(dev.pool, dev.stream, d_tmp, self.mont_one.ctypes.data, d_t + 32 * j, n)
            B.call("ff_poly_evaluate", cfg)
        # canonical t = mont_mul(1, t R)
        B.call("ff_ax", B.ptr(self.raw_one), ctypes.c_void_p(d_t), ctypes.c_void_p(d_t), ctypes.c_uint(m), dev.stream)
        d_p = dev.malloc(64 * m)
        B.call("bc_amd_generator_multiples", ctypes.c_void_p(d_t), ctypes.c_void_p(d_p), ctypes.c_uint(m), dev.stream)
        dev.sync()
        out = dev.download(d_p, (m, 16))
        for p in (d_t, d_tmp, d_p):
            dev.free(p)
        return out

    def folds(self, res_ptrs):
        """On-device fold of per-bit results to affine points."""
        dev = self.dev
        m = len(res_ptrs)
        arr = (ctypes.c_void_p * m)(*res_ptrs)
        d_p = dev.malloc(64 * m)
        B.call("bc_amd_msm_fold", arr, ctypes.c_void_p(d_p), ctypes.c_uint(m), dev.stream)
        dev.sync()
        out = dev.download(d_p, (m, 16))
        dev.free(d_p)
        return out


def msm_call(lab, d_bases, d_scalars, d_results, L):
    cfg = B.msm_configuration()
    cfg.mem_pool, cfg.stream = lab.dev.pool, lab.dev.stream
    cfg.bases, cfg.scalars, cfg.results, cfg.log_scalars_count = d_bases, d_scalars, d_results, L
    B.call("msm_execute_async", cfg)


def ntt_call(lab, d_in, d_out, L, bitrev=False, inverse=False, log_ext=0, coset=0):
    cfg = B.ntt_configuration()
    cfg.mem_pool, cfg.stream = lab.dev.pool, lab.dev.stream
    cfg.inputs, cfg.outputs, cfg.log_values_count = d_in, d_out, L
    cfg.bit_reversed_inputs, cfg.inverse = bitrev, inverse
    cfg.log_extension_degree, cfg.coset_index = log_ext, coset
    B.call("ntt_execute_async", cfg)


def pn_call(lab, d_idx, d_ks, d_out, cols, log_rows):
    cfg = B.generate_permutation_polynomials_configuration(lab.dev.pool, lab.dev.stream, d_idx, d_ks, d_out, cols, log_rows)
    B.call("pn_generate_permutation_polynomials", cfg)


# Fixed-base tables (bc_amd_msm_fixed_base_table) for MSM jobs over resident
# bases: three widths cover 2^12..2^22; the width per job size was the fastest
# of a per-size sweep of batches of equal jobs on MI355X, restricted to them.
FIXED_BASE_WIDTHS = (14, 17, 20)


def fixed_base_width(log_n):
    return 14 if log_n <= 15 else (17 if log_n <= 19 else 20)


def fixed_base_tables(lab, d_bases, log_count):
    """One table per width of FIXED_BASE_WIDTHS over 2^log_count bases: {c: device pointer}, seconds, bytes."""
    dev = lab.dev
    tables, nbytes = {}, 0
    t0 = time.perf_counter()
    for c in FIXED_BASE_WIDTHS:
        size = ((253 // c) + 1) * (64 << log_count)
        tables[c] = dev.malloc(size)
        nbytes += size
        B.call("bc_amd_msm_fixed_base_table", ctypes.c_void_p(d_bases), ctypes.c_uint(log_count), ctypes.c_uint(c), ctypes.c_void_p(tables[c]),
               dev.stream)
    dev.sync()
    return tables, time.perf_counter() - t0, nbytes


def pn_indexes(seed, cols, log_rows):
    """Variable indexes forming random cycles of length 1..64 over the cells."""
    cells = cols << log_rows
    rng = np.random.RandomState(seed & 0xFFFFFFFF)
    perm = rng.permutation(cells)
    lengths = rng.randint(1, 65, size=cells // 16 + 64)
    ends = np.cumsum(lengths)
    var_of_pos = np.searchsorted(ends, np.arange(cells), side="right")
    idx = np.zeros(cells, dtype=np.uint32)
    idx[perm] = var_of_pos.astype(np.uint32)
    return idx


def pct(v, q):
    return float(np.percentile(np.asarray(v), q)) if len(v) else None


# ---------------------------------------------------------------- autotuning
def autotune(lab, log_lo=12, log_hi=24):
    dev = lab.dev
    t0 = time.time()
    n = 1 << log_hi
    d_b, d_s = lab.bases(0xBC000010, n)
    d_k = lab.upload_mont(msm_scalars_canonical(0xBC000011, n))
    B.call("bc_amd_msm_autotune", ctypes.c_void_p(d_b), ctypes.c_void_p(d_k), ctypes.c_uint(log_lo), ctypes.c_uint(log_hi), dev.stream)
    scratch = dev.malloc(32 << 24)
    B.call("bc_amd_ntt_autotune", ctypes.c_void_p(scratch), ctypes.c_uint(13), ctypes.c_uint(24), dev.stream)
    for p in (d_b, d_s, d_k, scratch):
        dev.free(p)
    table = {}
    for L in range(log_lo, log_hi + 1):
        t = B.bc_amd_msm_tuning()
        B.call("bc_amd_msm_get_tuning", ctypes.c_uint(L), ctypes.byref(t))
        table["msm_%d" % L] = [t.window_bits, t.piece]
    for L in range(13, 25):
        f = ctypes.c_uint()
        B.call("bc_amd_ntt_get_split", ctypes.c_uint(L), ctypes.byref(f))
        table["ntt_%d" % L] = f.value
    log("autotune: %.1fs %s" % (time.time() - t0, table))
    return table


# ---------------------------------------------------------------- headline: 2^24 MSM + NTT
def headline_ntt(lab, L, steps, warmup):
    dev = lab.dev
    n = 1 << L
    x = uniform_fr(0xBC000003, n)
    d_x = lab.upload_mont(x)
    d_y = dev.malloc(n * 32)
    for _ in range(warmup):
        ntt_call(lab, d_x, d_y, L)
    dev.sync()
    ms = lab.time_ms(lambda: ntt_call(lab, d_x, d_y, L), steps)
    inv_ms = lab.time_ms(lambda: ntt_call(lab, d_y, d_y, L, bitrev=True, inverse=True), 1)
    # forward then inverse must give the input back (the inverse above ran in place on d_y)
    ok = bool(np.array_equal(dev.download(d_y, (n, 8)), dev.download(d_x, (n, 8))))
    lde_ms = lab.time_ms(lambda: [ntt_call(lab, d_x, d_y, L, log_ext=2, coset=c) for c in range(4)], max(1, steps // 2))
    dev.free(d_x)
    dev.free(d_y)
    return ms, inv_ms, lde_ms, ok


def headline_msm(lab, L, steps, warmup):
    dev = lab.dev
    n = 1 << L
    t0 = time.time()
    d_b, d_s = lab.bases(0xBC00002A, n)
    d_k = lab.upload_mont(msm_scalars_canonical(0xBC00002B, n))
    log("msm 2^%d: inputs ready in %.1fs" % (L, time.time() - t0))
    d_r = dev.malloc(254 * 96)
    for _ in range(warmup):
        msm_call(lab, d_b, d_k, d_r, L)
    dev.sync()
    ms = lab.time_ms(lambda: msm_call(lab, d_b, d_k, d_r, L), steps)
    got = lab.folds([d_r])[0]
    exp = lab.expected_points(d_s, [d_k], [n])[0]
    ok = bool(np.array_equal(got, exp))
    # the same MSM with the autotuned window width through the batch entry point
    job = B.msm_job(d_b, d_k, d_r, L)
    arr = B.array_of(B.bc_amd_msm_job, [job])
    tuned_ms = lab.time_ms(lambda: B.call("bc_amd_msm_batch_async", arr, ctypes.c_uint(1), dev.pool, dev.stream), steps)
    tuned_ok = bool(np.array_equal(lab.folds([d_r])[0], exp))
    for p in (d_b, d_s, d_k, d_r):
        dev.free(p)
    return ms, ok, tuned_ms, tuned_ok


# ---------------------------------------------------------------- PN (the served mix's 4 x 2^20 job)
def headline_pn(lab, steps, warmup, cols=4, log_rows=20):
    """Permutation polynomials for 4 x 2^20 cells (random cycles 1..64) through
    the plain C ABI; the roofline counts 36 B per cell (a 4 B index read and a
    32 B value written)."""
    dev = lab.dev
    cells = cols << log_rows
    d_idx = dev.upload(pn_indexes(0xBC00000F, cols, log_rows))
    d_ks = lab.upload_mont(uniform_fr(0xBC000010, cols))
    d_out = dev.malloc(32 * cells)
    for _ in range(warmup):
        pn_call(lab, d_idx, d_ks, d_out, cols, log_rows)
    dev.sync()
    ms = lab.time_ms(lambda: pn_call(lab, d_idx, d_ks, d_out, cols, log_rows), steps)
    for p in (d_idx, d_ks, d_out):
        dev.free(p)
    achieved = 36.0 * cells / (ms * 1e-3) / 1e9
    return {"workload": "PN %d x 2^%d (random cycles 1..64)" % (cols, log_rows), "ms": round(ms, 3),
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "algorithmic_bytes": 36 * cells, "kernel": "whole pn_generate_permutation_polynomials call"}}


# ---------------------------------------------------------------- batched (64 MSM + 32 NTT)
def batched(lab, reps):
    dev = lab.dev
    rng = np.random.RandomState(0xBC000004 & 0xFFFFFFFF)
    msm_logs = [int(v) for v in rng.randint(12, 21, 64)]
    ntt_logs = [int(v) for v in rng.randint(10, 23, 32)]
    t0 = time.time()
    d_b, d_s = lab.bases(0xBC000002, 1 << 20)
    d_k, r_plain, r_batch = [], [], []
    for j, L in enumerate(msm_logs):
        d_k.append(lab.upload_mont(msm_scalars_canonical(0xBC000040 + j, 1 << L)))
        r_plain.append(dev.malloc(254 * 96))
        r_batch.append(dev.malloc(254 * 96))
    x_pool = lab.upload_mont(uniform_fr(0xBC000004, 1 << 23))
    x_in, y_plain, y_batch = [], [], []
    for j, L in enumerate(ntt_logs):
        off = (j * 1000003) % ((1 << 23) - (1 << L) + 1)
        x_in.append(x_pool + 32 * off)
        y_plain.append(dev.malloc(32 << L))
        y_batch.append(dev.malloc(32 << L))
    dev.sync()
    log("batched: inputs ready in %.1fs (MSM logs %s, NTT logs %s)" % (time.time() - t0, msm_logs, ntt_logs))

    def one_at_a_time():
        for j, L in enumerate(msm_logs):
            msm_call(lab, d_b, d_k[j], r_plain[j], L)
            dev.sync()
        for j, L in enumerate(ntt_logs):
            ntt_call(lab, x_in[j], y_plain[j], L)
            dev.sync()

    for _ in range(2):  # warm-up (pool growth, twiddles, first launches)
        one_at_a_time()
    plain_s = []
    for _ in range(reps):
        t = time.perf_counter()
        one_at_a_time()
        plain_s.append(time.perf_counter() - t)
    plain_s = float(np.median(plain_s))

    msm_jobs = [B.msm_job(d_b, d_k[j], r_batch[j], L) for j, L in enumerate(msm_logs)]
    ntt_jobs = [B.ntt_job(x_in[j], y_batch[j], L) for j, L in enumerate(ntt_logs)]
    marr, narr = B.array_of(B.bc_amd_msm_job, msm_jobs), B.array_of(B.bc_amd_ntt_job, ntt_jobs)

    def batch_calls():
        B.call("bc_amd_msm_batch_async", marr, ctypes.c_uint(len(msm_jobs)), dev.pool, dev.stream)
        B.call("bc_amd_ntt_batch_async", narr, ctypes.c_uint(len(ntt_jobs)), dev.pool, dev.stream)

    batch_calls()
    dev.sync()
    batch_ms = lab.time_ms(batch_calls, reps)
    plan = B.bc_amd_plan()
    t = time.perf_counter()
    B.call("bc_amd_plan_create", ctypes.byref(plan), marr, ctypes.c_uint(len(msm_jobs)), narr, ctypes.c_uint(len(ntt_jobs)))
    plan_create_s = time.perf_counter() - t
    ws = ctypes.c_size_t()
    B.call("bc_amd_plan_workspace_bytes", plan, ctypes.byref(ws))
    B.call("bc_amd_plan_launch", plan, dev.stream)
    dev.sync()
    plan_ms = lab.time_ms(lambda: B.call("bc_amd_plan_launch", plan, dev.stream), reps)
    t = time.perf_counter()
    for _ in range(reps):
        B.call("bc_amd_plan_launch", plan, dev.stream)
    dev.sync()
    plan_wall_s = (time.perf_counter() - t) / reps
    B.call("bc_amd_plan_destroy", plan)

    # the same jobs with the MSMs over fixed-base tables of the shared bases
    tables, table_s, table_bytes = fixed_base_tables(lab, d_b, 20)
    r_fixed = [dev.malloc(254 * 96) for _ in msm_logs]
    fb_jobs = [B.msm_job(tables[fixed_base_width(L)], d_k[j], r_fixed[j], L, fixed_base_width(L), 0, 20) for j, L in enumerate(msm_logs)]
    fbarr = B.array_of(B.bc_amd_msm_job, fb_jobs)
    plan_fb = B.bc_amd_plan()
    B.call("bc_amd_plan_create", ctypes.byref(plan_fb), fbarr, ctypes.c_uint(len(fb_jobs)), narr, ctypes.c_uint(len(ntt_jobs)))
    B.call("bc_amd_plan_launch", plan_fb, dev.stream)
    dev.sync()
    t = time.perf_counter()
    for _ in range(reps):
        B.call("bc_amd_plan_launch", plan_fb, dev.stream)
    dev.sync()
    plan_fb_wall_s = (time.perf_counter() - t) / reps
    B.call("bc_amd_plan_destroy", plan_fb)

    # parity: NTT bytes identical to the plain calls, MSM folds identical and equal to the closed form
    ntt_ok = all(np.array_equal(dev.download(y_plain[j], (1 << L, 8)), dev.download(y_batch[j], (1 << L, 8))) for j, L in enumerate(ntt_logs))
    f_plain, f_batch = lab.folds(r_plain), lab.folds(r_batch)
    exp = lab.expected_points(d_s, d_k, [1 << L for L in msm_logs])
    msm_ok = bool(np.array_equal(f_plain, f_batch) and np.array_equal(f_batch, exp))
    fb_ok = bool(np.array_equal(lab.folds(r_fixed), exp))
    for p in [d_b, d_s, x_pool] + d_k + r_plain + r_batch + y_plain + y_batch + r_fixed + list(tables.values()):
        dev.free(p)
    jobs = len(msm_logs) + len(ntt_logs)
    return {
        "workload": "64 MSMs 2^12..2^20 (shared 2^20 bases) + 32 forward NTTs 2^10..2^22",
        "jobs": jobs,
        "one_at_a_time_jobs_per_s": round(jobs / plain_s, 1),
        "one_at_a_time_ms": round(plain_s * 1e3, 2),
        "plan_wall_ms": round(plan_fb_wall_s * 1e3, 2),
        "plan_jobs_per_s": round(jobs / plan_fb_wall_s, 1),
        "speedup_vs_one_at_a_time": round(plain_s / plan_fb_wall_s, 2),
        "fixed_base": {"widths": list(FIXED_BASE_WIDTHS), "tables_s": round(table_s, 3), "tables_gib": round(table_bytes / 2**30, 2),
                       "note": "the plan's MSMs run over fixed-base tables of the resident bases (built once, outside the timed region)"},
        "bases_jobs": {"batch_calls_ms": round(batch_ms, 2), "plan_ms": round(plan_ms, 2), "plan_wall_ms": round(plan_wall_s * 1e3, 2),
                       "plan_jobs_per_s": round(jobs / plan_wall_s, 1), "speedup_vs_one_at_a_time": round(plain_s / plan_wall_s, 2),
                       "note": "the same plan with the MSMs over the bases themselves"},
        "plan_create_s": round(plan_create_s, 2),
        "plan_workspace_gib": round(ws.value / 2**30, 2),
        "parity": {"ntt_bytes_equal_plain": bool(ntt_ok), "msm_folds_equal_plain_and_closed_form": msm_ok,
                   "fixed_base_msm_folds_equal_closed_form": fb_ok},
    }


# ---------------------------------------------------------------- served mix
class ServedJobs:
    def __init__(self, lab, n_jobs, seed=0xBC000005):
        dev = lab.dev
        self.lab = lab
        rng = np.random.RandomState(seed & 0xFFFFFFFF)
        t0 = time.time()
        self.d_b, self.d_s = lab.bases(seed, 1 << 22)
        self.k_pool = lab.upload_mont(msm_scalars_canonical(seed + 1, (1 << 22) + (1 << 21)))
        self.x_pool = lab.upload_mont(uniform_fr(seed + 2, 1 << 23))
        self.pn_idx = [dev.upload(pn_indexes(seed + 3 + t, 4, 20)) for t in range(2)]
        self.pn_ks = lab.upload_mont(uniform_fr(seed + 5, 4))
        self.jobs = []  # (kind, params, output ptr)
        for j in range(n_jobs):
            u = rng.randint(100)
            if u < 50:
                L = int(rng.randint(14, 23))
                off = int(rng.randint(0, 1 << 21))
                self.jobs.append(("msm", (L, self.k_pool + 32 * off, off), dev.malloc(254 * 96)))
            elif u < 85:
                L = int(rng.randint(12, 23))
                off = int(rng.randint(0, (1 << 23) - (1 << L) + 1))
                self.jobs.append(("ntt", (L, self.x_pool + 32 * off), dev.malloc(32 << L)))
            else:
                self.jobs.append(("pn", (self.pn_idx[j % 2], 4, 20), dev.malloc(32 * (4 << 20))))
        dev.sync()
        log("served: %d jobs ready in %.1fs (%d MSM, %d NTT, %d PN)" % (n_jobs, time.time() - t0, *[sum(1 for k in self.jobs if k[0] == t) for t in
                                                                                                       ("msm", "ntt", "pn")]))

    def plain(self, j):
        kind, prm, out = self.jobs[j]
        if kind == "msm":
            msm_call(self.lab, self.d_b, prm[1], out, prm[0])
        elif kind == "ntt":
            ntt_call(self.lab, prm[1], out, prm[0])
        else:
            pn_call(self.lab, prm[0], self.pn_ks, out, prm[1], prm[2])

    def job(self, j, tables=None):
        kind, prm, out = self.jobs[j]
        job = B.bc_amd_job()
        if kind == "msm":
            job.kind = B.BC_AMD_JOB_MSM
            if tables:
                c = fixed_base_width(prm[0])
                job.msm = B.msm_job(tables[c], prm[1], out, prm[0], c, 0, 22)
            else:
                job.msm = B.msm_job(self.d_b, prm[1], out, prm[0])
        elif kind == "ntt":
            job.kind = B.BC_AMD_JOB_NTT
            job.ntt = B.ntt_job(prm[1], out, prm[0])
        else:
            job.kind = B.BC_AMD_JOB_PN
            job.pn = B.bc_amd_pn_job(prm[0], self.pn_ks, out, prm[1], prm[2])
        return job

    def clear(self):
        for kind, prm, out in self.jobs:
            n = 254 * 96 if kind == "msm" else (32 << prm[0] if kind == "ntt" else 32 * (4 << 20))
            B.call("bc_memset", ctypes.c_void_p(out), ctypes.c_int(0), ctypes.c_size_t(n))

    def fingerprints(self):
        """NTT/PN: digest of the output bytes; MSM: the folded affine point."""
        dev = self.lab.dev
        msm_idx = [j for j, (k, _, _) in enumerate(self.jobs) if k == "msm"]
        folds = self.lab.folds([self.jobs[j][2] for j in msm_idx]) if msm_idx else []
        fp = {}
        for j, f in zip(msm_idx, folds):
            fp[j] = f.tobytes()
        for j, (kind, prm, out) in enumerate(self.jobs):
            if kind == "ntt":
                fp[j] = hashlib.sha1(dev.download(out, (1 << prm[0], 8)).tobytes()).hexdigest()
            elif kind == "pn":
                fp[j] = hashlib.sha1(dev.download(out, (4 << 20, 8)).tobytes()).hexdigest()
        return fp

    def msm_closed_form_ok(self, fp):
        msm_idx = [j for j, (k, _, _) in enumerate(self.jobs) if k == "msm"]
        if not msm_idx:
            return True
        # s prefix [0, n) pairs with the scalar slice [off, off + n)
        exp = self.lab.expected_points(self.d_s, [self.jobs[j][1][1] for j in msm_idx], [1 << self.jobs[j][1][0] for j in msm_idx])
        return all(fp[j] == e.tobytes() for j, e in zip(msm_idx, exp))

    def free(self):
        dev = self.lab.dev
        for p in [self.d_b, self.d_s, self.k_pool, self.x_pool, self.pn_ks] + self.pn_idx + [o for _, _, o in self.jobs]:
            dev.free(p)


def run_served(srv, sj, arrivals=None, tables=None):
    """Submits every job (at `arrivals` seconds from start, or at once); returns stats per job."""
    n = len(sj.jobs)
    tickets = [None] * n
    stats = [None] * n
    jobs = [sj.job(j, tables) for j in range(n)]

    def waiter():
        for j in range(n):
            while tickets[j] is None:
                time.sleep(50e-6)
            stats[j] = srv.wait(tickets[j])

    th = threading.Thread(target=waiter)
    th.start()
    t0 = time.perf_counter()
    for j in range(n):
        if arrivals is not None:
            delay = t0 + arrivals[j] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
        tickets[j] = srv.submit(jobs[j])
    th.join()
    return stats


def served(lab, n_jobs, budget_gib, slots):
    dev = lab.dev
    sj = ServedJobs(lab, n_jobs)
    n = len(sj.jobs)
    # one at a time through the plain C ABI (sync after each), also the reference outputs
    for j in range(min(n, 8)):
        sj.plain(j)
        dev.sync()
    sj.clear()
    t = time.perf_counter()
    plain_done = []
    for j in range(n):
        sj.plain(j)
        dev.sync()
        plain_done.append(time.perf_counter() - t)
    plain_s = time.perf_counter() - t
    fp_plain = sj.fingerprints()
    closed_ok = sj.msm_closed_form_ok(fp_plain)

    tables, table_s, table_bytes = fixed_base_tables(lab, sj.d_b, 22)
    srv = B.Server(hbm_budget_bytes=int(budget_gib * 2**30), max_in_flight=slots)
    try:
        # capacity with the MSMs over the bases themselves
        run_served(srv, sj)  # warm-up (pool growth)
        sj.clear()
        st = run_served(srv, sj)
        cap_plain_jobs = n / ((max(s.done_ns for s in st) - min(s.submit_ns for s in st)) / 1e9)
        mismatches_plain_jobs = [j for j, f in sj.fingerprints().items() if f != fp_plain[j]]
        # the served form: MSMs over fixed-base tables of the resident bases
        run_served(srv, sj, tables=tables)
        sj.clear()
        st = run_served(srv, sj, tables=tables)
        cap_s = (max(s.done_ns for s in st) - min(s.submit_ns for s in st)) / 1e9
        t0 = min(s.submit_ns for s in st)
        batches = sorted({(s.start_ns, s.enqueued_ns, s.done_ns, s.batch_jobs, s.slot) for s in st})
        timeline = [[round((a - t0) / 1e6, 2), round((q - t0) / 1e6, 2), round((d - t0) / 1e6, 2), n, k] for a, q, d, n, k in batches]
        submit_span = (max(s.submit_ns for s in st) - t0) / 1e6
        fp_served = sj.fingerprints()
        mismatches = [j for j in range(n) if fp_served[j] != fp_plain[j]]
        capacity = n / cap_s
        loads = {}
        rng = np.random.RandomState(0xBC000006)
        trace07 = None
        for frac in (0.5, 0.7, 0.9):
            lam = frac * capacity
            arrivals = np.cumsum(rng.exponential(1.0 / lam, n))
            if frac == 0.7:
                trace07 = arrivals
            st = run_served(srv, sj, arrivals, tables=tables)
            lat = [(s.done_ns - s.submit_ns) / 1e6 for s in st]
            span = (max(s.done_ns for s in st) - min(s.submit_ns for s in st)) / 1e9
            loads["%.1f" % frac] = {"rate_jobs_per_s": round(lam, 1), "p50_ms": round(pct(lat, 50), 2), "p99_ms": round(pct(lat, 99), 2),
                                    "jobs_per_s": round(n / span, 1), "mean_batch": round(float(np.mean([s.batch_jobs for s in st])), 2)}
            log("served at %.1f x capacity: %s" % (frac, loads["%.1f" % frac]))
    finally:
        srv.close()
    # the one-at-a-time FIFO under the same 0.7 x capacity arrivals
    lat_plain = []
    t0 = time.perf_counter()
    for j in range(n):
        now = time.perf_counter() - t0
        if now < trace07[j]:
            time.sleep(trace07[j] - now)
        sj.plain(j)
        dev.sync()
        lat_plain.append((time.perf_counter() - t0 - trace07[j]) * 1e3)
    sj.free()
    for p in tables.values():
        dev.free(p)
    return {
        "workload": "Poisson mix: 50% MSM 2^14..2^22, 35% NTT 2^12..2^22, 15% PN 4x2^20 (random cycles 1..64)",
        "jobs": n,
        "server": {"hbm_budget_gib": budget_gib, "slots": slots},
        "one_at_a_time_jobs_per_s": round(n / plain_s, 1),
        "served_capacity_jobs_per_s": round(capacity, 1),
        "speedup_vs_one_at_a_time": round(capacity / (n / plain_s), 2),
        "fixed_base": {"widths": list(FIXED_BASE_WIDTHS), "tables_s": round(table_s, 3), "tables_gib": round(table_bytes / 2**30, 2),
                       "note": "served MSMs run over fixed-base tables of the resident 2^22 bases (built once)"},
        "bases_jobs": {"served_capacity_jobs_per_s": round(cap_plain_jobs, 1), "speedup_vs_one_at_a_time": round(cap_plain_jobs / (n / plain_s), 2),
                       "note": "capacity with the MSMs over the bases themselves"},
        "capacity_batches_ms": timeline,
        "capacity_submit_span_ms": round(submit_span, 2),
        "latency_under_poisson": loads,
        "one_at_a_time_fifo_at_0.7": {"p50_ms": round(pct(lat_plain, 50), 2), "p99_ms": round(pct(lat_plain, 99), 2)},
        "parity": {"served_equal_plain": not mismatches and not mismatches_plain_jobs, "mismatched_jobs": (mismatches + mismatches_plain_jobs)[:8],
                   "msm_closed_form": bool(closed_ok)},
    }


# ---------------------------------------------------------------- CPU baseline (the oracle, host cores)
def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def median_time(fn, reps):
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def cpu_baseline():
    """The C oracle (the reference's host_math arithmetic restated) on this
    host's cores, on bounded samples of the configurations: C1 (NTT 2^16) on 1
    thread and all threads, median of 5; NTT 2^20 and the C2 MSM 2^20 on all
    threads, median of 3, and on one thread (NTT 2^20 once, MSM 2^16 median of
    3); the C5 PN job (4 x 2^20, serial restatement) median of 3.  `value` is
    the headline step extrapolated from the all-thread 2^20 samples by n log n
    (NTT) and n W (MSM)."""
    from oracle import oracle as O

    all_threads = O.get_threads()
    out = {"cpu_model": cpu_model(), "nproc": os.cpu_count(), "threads_all": all_threads}
    x16 = O.random_fr(0xBC000001, 1 << 16)
    x20 = O.random_fr(0xBC000003, 1 << 20)
    n = 1 << 20
    bases = O.gen_bases_progression(3, 5, n)
    k = O.msm_scalars(0xBC000002, n)
    for threads in (1, all_threads):
        O.set_threads(threads)
        tag = "1t" if threads == 1 else "all"
        out["c1_ntt_2e16_ms_" + tag] = round(median_time(lambda: O.ntt(x16, 16), 5) * 1e3, 2)
        if threads == 1:
            out["ntt_2e20_ms_1t"] = round(median_time(lambda: O.ntt(x20, 20), 1) * 1e3, 1)
            out["msm_2e16_ms_1t"] = round(median_time(lambda: O.msm(bases[: 1 << 16], k[: 1 << 16], 16), 3) * 1e3, 1)
        else:
            out["ntt_2e20_ms_all"] = round(median_time(lambda: O.ntt(x20, 20), 3) * 1e3, 1)
            out["c2_msm_2e20_ms_all"] = round(median_time(lambda: O.msm(bases, k, 20), 3) * 1e3, 1)
    O.set_threads(all_threads)
    idx = pn_indexes(0xBC00000F, 4, 20)
    ks = O.random_fr(0xBC000010, 4)
    out["c5_pn_4x2e20_ms_1t"] = round(median_time(lambda: O.pn(idx, ks, 4, 20), 3) * 1e3, 1)
    c20, c24 = O.msm_window_bits(20), O.msm_window_bits(MSM_METRIC_LOG)
    w20, w24 = 253 // c20 + 1, 253 // c24 + 1
    ntt24 = out["ntt_2e20_ms_all"] * (24 * 2**24) / (20 * 2**20)
    msm24 = out["c2_msm_2e20_ms_all"] * (2**24 * w24) / (2**20 * w20)
    out.update({
        "value": round(ntt24 + msm24, 1),
        "unit": "ms per step (MSM 2^24 + NTT 2^24), extrapolated from the 2^20 samples",
        "cores": all_threads,
        "kind": "port",
        "sample": "C oracle on %d threads of %s: NTT 2^20 %.0f ms and MSM 2^20 %.0f ms (medians of 3), scaled by n log n and n W; "
        "1-thread and C1/C5 timings alongside" % (all_threads, out["cpu_model"], out["ntt_2e20_ms_all"], out["c2_msm_2e20_ms_all"]),
    })
    return out


def profiled_kernel_ms(name):
    """Average duration of `name` in the plain 2^24 MSM call, from the committed rocprofv3 kernel-trace summary."""
    path = os.path.join(ROOT, "profiles", "msm_ntt_2e24_kernel_trace.md")
    try:
        for line in open(path):
            if line.startswith("avg_ms %s:" % name):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def hbm_copy_ceiling(lab, nbytes=1 << 30, reps=5):
    """Device-to-device copy of 1 GiB through bc_memcpy_async (the runtime's
    copy kernel): read + written bytes per second, the achievable HBM rate."""
    dev = lab.dev
    a, b = dev.malloc(nbytes), dev.malloc(nbytes)
    cp = lambda: B.call("bc_memcpy_async", ctypes.c_void_p(b), ctypes.c_void_p(a), ctypes.c_size_t(nbytes), dev.stream)
    cp()
    dev.sync()
    ms = lab.time_ms(cp, reps)
    dev.free(a)
    dev.free(b)
    return {"gbs": round(2.0 * nbytes / (ms * 1e-3) / 1e9, 1), "ms": round(ms, 3), "bytes_each_way": nbytes}


def pmc_traffic():
    """HBM bytes per ntt_pass_kernel launch from the committed rocprofv3 counter summary, if present."""
    path = os.path.join(ROOT, "profiles", "ntt_pass_kernel_2e24_counters.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--log-msm", type=int, default=MSM_METRIC_LOG)
    ap.add_argument("--log-ntt", type=int, default=24)
    ap.add_argument("--parts", default="tune,ntt,msm,pn,batched,served,cpu",
                    help="comma list of: tune, ntt, msm, batched, served, cpu, copy (a 2^24-element copy kernel, for counter scaling)")
    ap.add_argument("--served-jobs", type=int, default=120)
    ap.add_argument("--served-budget-gib", type=float, default=160.0)
    ap.add_argument("--served-slots", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the hot path is built and measured on a single MI355X; run it with --gpus 1", file=sys.stderr)
        sys.exit(2)
    parts = set(args.parts.split(","))
    lab = Lab()
    log("device ready")
    out = {
        "metric": "ms per BN254 MSM/NTT at 2^24; batched jobs/s + p99 vs one-at-a-time calls",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 limbs (exact Fq/Fr Montgomery arithmetic)",
        "data": "synthetic, seeded (splitmix64)",
        "config": {"workload": "BN254 G1 MSM 2^%d + Fr NTT 2^%d (forward, natural in, bit-reversed out)" % (args.log_msm, args.log_ntt),
                   "msm_log_n": args.log_msm, "ntt_log_n": args.log_ntt},
    }
    if "copy" in parts:
        n = 1 << 24
        src, dst = lab.dev.malloc(32 * n), lab.dev.malloc(32 * n)
        for _ in range(3):
            B.call("pn_distribute_values", ctypes.c_void_p(src), ctypes.c_void_p(dst), ctypes.c_uint(n), ctypes.c_uint(1), lab.dev.stream)
        lab.dev.sync()
        log("copy kernel: 3 launches of pn_distribute_values, 2^24 elements (512 MiB read + 512 MiB written each)")
    if "ntt" in parts:
        out["hbm_copy_ceiling"] = hbm_copy_ceiling(lab)
        log("hbm copy ceiling: %s" % out["hbm_copy_ceiling"])
    if "tune" in parts:
        out["tuning"] = autotune(lab)
    if "ntt" in parts:
        ntt_ms, inv_ms, lde_ms, ntt_ok = headline_ntt(lab, args.log_ntt, args.steps, args.warmup)
        log("ntt 2^%d: %.3f ms, inverse %.3f ms, lde x4 %.3f ms, round trip ok=%s" % (args.log_ntt, ntt_ms, inv_ms, lde_ms, ntt_ok))
        n = 1 << args.log_ntt
        launches = (args.log_ntt + 9) // 10  # passes of at most 10 stages
        per_launch_ms = ntt_ms / launches
        # algorithmic bytes: 64 n per NTT call (each 32 B element read and written
        # once), i.e. 64 n / passes per launch of the pass kernel
        bytes_per_launch = 64.0 * n / launches
        achieved = bytes_per_launch / (per_launch_ms * 1e-3) / 1e9
        out.update({"ntt_ms": round(ntt_ms, 3), "ntt_inverse_ms": round(inv_ms, 3), "lde_4_cosets_ms": round(lde_ms, 3),
                    "ntt_call_gbs": round(64.0 * n / (ntt_ms * 1e-3) / 1e9, 1)})
        traffic = pmc_traffic()
        out["roofline"] = {"bound": "hbm", "kernel": "ntt_pass_kernel (2^%d, %d launches per NTT)" % (args.log_ntt, launches),
                           "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                           "traffic": traffic, "algorithmic_bytes_per_launch": int(bytes_per_launch), "launch_ms": round(per_launch_ms, 4)}
        if traffic:
            # counter-measured HBM bytes per launch (profiles/ summary) over this run's launch time
            per_pass = traffic / (per_launch_ms * 1e-3) / 1e9
            out["hbm_per_pass"] = {"gbs": round(per_pass, 1), "frac": round(per_pass / HBM_PEAK_GBS, 4),
                                   "source": "rocprofv3 FETCH_SIZE/WRITE_SIZE per ntt_pass_kernel launch, scaled by a copy kernel"}
        out.setdefault("parity", {})["ntt_round_trip"] = ntt_ok
    if "msm" in parts:
        msm_ms, msm_ok, tuned_ms, tuned_ok = headline_msm(lab, args.log_msm, args.steps, args.warmup)
        log("msm 2^%d: %.3f ms (closed form ok=%s), autotuned width %.3f ms (ok=%s)" % (args.log_msm, msm_ms, msm_ok, tuned_ms, tuned_ok))
        out.update({"msm_ms": round(msm_ms, 3), "msm_autotuned_ms": round(tuned_ms, 3)})
        c = 20 if args.log_msm == 24 else None
        if c:
            W = 253 // c + 1
            products = (1 << args.log_msm) * W * 10 * 136
            tps = products / (msm_ms * 1e-3) / 1e12
            out["roofline_msm"] = {"bound": "valu", "kernel": "whole msm_execute_async call", "achieved": round(tps, 2), "peak": VALU_ISSUE_PEAK_T,
                                   "unit": "T 32x32-bit products/s", "frac": round(tps / VALU_ISSUE_PEAK_T, 4),
                                   "mad_probe_t": VALU_MAD_PROBE_T, "frac_of_probe": round(tps / VALU_MAD_PROBE_T, 4),
                                   "work": "2^24 x %d windows x 10 Fq multiplies x 136 products = %.3g" % (W, products)}
            acc_ms = profiled_kernel_ms("accumulate_kernel")
            if acc_ms:
                at = products / (acc_ms * 1e-3) / 1e12
                out["roofline_msm"]["accumulate_kernel"] = {"ms": acc_ms, "achieved": round(at, 2), "frac": round(at / VALU_ISSUE_PEAK_T, 4),
                                                            "source": "rocprofv3 kernel trace summary in profiles/"}
        out.setdefault("parity", {}).update({"msm_closed_form": msm_ok, "msm_autotuned_closed_form": tuned_ok})
    if "pn" in parts:
        out["pn"] = headline_pn(lab, max(1, args.steps), args.warmup)
        log("pn: %s" % out["pn"])
    if "ntt" in parts and "msm" in parts:
        step = out["msm_ms"] + out["ntt_ms"]
        out.update({"value": round(step, 3), "unit": "ms per step (one 2^24 G1 MSM + one 2^24 Fr NTT)", "ms_per_step": round(step, 3)})
    if "batched" in parts:
        out["batched"] = batched(lab, max(1, args.steps))
        log("batched: %s" % out["batched"])
    if "served" in parts:
        out["served"] = served(lab, args.served_jobs, args.served_budget_gib, args.served_slots)
        log("served: %s" % out["served"])
    if "cpu" in parts:
        out["cpu_baseline"] = cpu_baseline()
        log("cpu baseline: %s" % out["cpu_baseline"])
    lab.dev.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
