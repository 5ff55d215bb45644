"""Benchmark of the prover hot path on one MI355X.

A step is one BN254 G1 MSM of 2^24 points plus one Fr NTT of 2^24 elements
(forward, natural -> bit-reversed), both called through the bellman-cuda C ABI
with inputs resident in HBM.  Inputs are synthetic and seeded: bases
[s_i]G for uniform s_i (generated on the device by a fixed-base comb), MSM
scalars 70% uniform / 20% < 2^64 / 10% in {0,1}, uniform NTT inputs.
Prints one JSON line.  The CPU baseline is the C oracle timed on a bounded
sample on the host.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "era-bellman-cuda_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

import bellman_amd as B  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class Timer:
    def __init__(self, dev):
        self.dev = dev
        self.start, self.stop = B.bc_event(), B.bc_event()
        B.call("bc_event_create", ctypes.byref(self.start), ctypes.c_bool(False), ctypes.c_bool(False))
        B.call("bc_event_create", ctypes.byref(self.stop), ctypes.c_bool(False), ctypes.c_bool(False))

    def time(self, fn, reps):
        B.call("bc_event_record", self.start, self.dev.stream)
        for _ in range(reps):
            fn()
        B.call("bc_event_record", self.stop, self.dev.stream)
        B.call("bc_event_synchronize", self.stop)
        ms = ctypes.c_float()
        B.call("bc_event_elapsed_time", ctypes.byref(ms), self.start, self.stop)
        return ms.value / reps


def msm_call(dev, d_bases, d_scalars, d_results, L):
    cfg = B.msm_configuration()
    cfg.mem_pool, cfg.stream = dev.pool, dev.stream
    cfg.bases, cfg.scalars, cfg.results, cfg.log_scalars_count = d_bases, d_scalars, d_results, L
    B.call("msm_execute_async", cfg)


def ntt_call(dev, d_in, d_out, L, bitrev=False, inverse=False, log_ext=0, coset=0):
    cfg = B.ntt_configuration()
    cfg.mem_pool, cfg.stream = dev.pool, dev.stream
    cfg.inputs, cfg.outputs, cfg.log_values_count = d_in, d_out, L
    cfg.bit_reversed_inputs, cfg.inverse = bitrev, inverse
    cfg.log_extension_degree, cfg.coset_index = log_ext, coset
    B.call("ntt_execute_async", cfg)


def msm_checked(dev, O, L, steps, warmup, timer):
    n = 1 << L
    t0 = time.time()
    s = O.random_fr(0xBC000002A, n, montgomery=False)
    k = O.msm_scalars(0xBC000002B, n)
    log("msm 2^%d: inputs generated on host in %.1fs" % (L, time.time() - t0))
    d_s = dev.upload(s)
    d_k = dev.upload(k)
    d_bases = dev.malloc(n * 64)
    B.call("bc_amd_generator_multiples", ctypes.c_void_p(d_s), ctypes.c_void_p(d_bases), ctypes.c_uint(n), dev.stream)
    d_res = dev.malloc(254 * 96)
    dev.sync()
    for _ in range(warmup):
        msm_call(dev, d_bases, d_k, d_res, L)
    dev.sync()
    ms = timer.time(lambda: msm_call(dev, d_bases, d_k, d_res, L), steps)
    res = dev.download(d_res, (254, 24))
    # size-independent check: fold == [sum k_i s_i] G, the sum taken on the device with ff ops
    d_sm = dev.malloc(n * 32)
    r2 = O.to_limbs([(1 << 512) % O.R])
    B.call("ff_ax", B.ptr(r2), ctypes.c_void_p(d_s), ctypes.c_void_p(d_sm), ctypes.c_uint(n), dev.stream)  # s -> Montgomery
    B.call("ff_x_mul_y", ctypes.c_void_p(d_sm), ctypes.c_void_p(d_k), ctypes.c_void_p(d_sm), ctypes.c_uint(n), dev.stream)
    one = O.to_mont([1])
    tsum = np.zeros((1, 8), dtype=np.uint32)
    B.call("ff_poly_evaluate", B.ff_poly_evaluate_configuration(dev.pool, dev.stream, d_sm, one.ctypes.data, tsum.ctypes.data, n))
    dev.sync()
    t = O.from_mont(tsum)[0]
    ok = bool(np.array_equal(O.to_affine(O.msm_fold(res))[0], O.mul_generator(t)))
    for p in (d_s, d_k, d_bases, d_res, d_sm):
        dev.free(p)
    return ms, ok


def ntt_checked(dev, O, L, steps, warmup, timer):
    n = 1 << L
    x = O.random_fr(0xBC000003, n)
    d_x = dev.upload(x)
    d_y = dev.malloc(n * 32)
    for _ in range(warmup):
        ntt_call(dev, d_x, d_y, L)
    dev.sync()
    ms = timer.time(lambda: ntt_call(dev, d_x, d_y, L), steps)
    lde_ms = timer.time(lambda: [ntt_call(dev, d_x, d_y, L, log_ext=2, coset=c) for c in range(4)], max(1, steps // 2))
    # round trip: inverse bit-reversed -> natural brings the input back
    ntt_call(dev, d_x, d_y, L)
    ntt_call(dev, d_y, d_y, L, bitrev=True, inverse=True)
    dev.sync()
    ok = bool(np.array_equal(dev.download(d_y, (n, 8)), x))
    dev.free(d_x)
    dev.free(d_y)
    return ms, lde_ms, ok


def cpu_baseline(O, L_ntt_sample=20, L_msm_sample=16):
    """The C oracle on the host: NTT 2^20 on one thread, MSM 2^16 (OpenMP over windows)."""
    x = O.random_fr(1, 1 << L_ntt_sample)
    t0 = time.time()
    O.ntt(x, L_ntt_sample)
    t_ntt = time.time() - t0
    n = 1 << L_msm_sample
    bases = O.gen_bases_progression(3, 5, n)
    k = O.msm_scalars(7, n)
    t0 = time.time()
    O.msm(bases, k, L_msm_sample)
    t_msm = time.time() - t0
    c16, c24 = O.msm_window_bits(L_msm_sample), O.msm_window_bits(24)
    w16, w24 = (253 // c16) + 1, (253 // c24) + 1
    # scale by the dominant work: n log n butterflies for the NTT, n * W mixed additions for the MSM
    ntt24 = t_ntt * (24 * 2**24) / (L_ntt_sample * 2**L_ntt_sample)
    msm24 = t_msm * (2**24 * w24) / (2**L_msm_sample * w16)
    return {
        "value": round((ntt24 + msm24) * 1000.0, 1),
        "unit": "ms per step (MSM 2^24 + NTT 2^24), extrapolated",
        "cores": os.cpu_count(),
        "kind": "port",
        "sample": "oracle NTT 2^%d on 1 thread (%.2fs) and MSM 2^%d with OpenMP over windows (%.2fs), scaled by n*log n and n*W"
        % (L_ntt_sample, t_ntt, L_msm_sample, t_msm),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--log-msm", type=int, default=24)
    ap.add_argument("--log-ntt", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the hot path is built and measured on a single MI355X (--gpus 1)", file=sys.stderr)
        sys.exit(2)
    from oracle import oracle as O

    B.call("ff_set_up", ctypes.c_uint(25), ctypes.c_uint(14))
    B.call("ntt_set_up")
    B.call("msm_set_up")
    dev = B.Device(0)
    timer = Timer(dev)
    log("device ready")
    ntt_ms, lde_ms, ntt_ok = ntt_checked(dev, O, args.log_ntt, args.steps, args.warmup, timer)
    log("ntt 2^%d: %.3f ms, lde x4: %.3f ms, round trip ok=%s" % (args.log_ntt, ntt_ms, lde_ms, ntt_ok))
    msm_ms, msm_ok = msm_checked(dev, O, args.log_msm, args.steps, args.warmup, timer)
    log("msm 2^%d: %.3f ms, closed form ok=%s" % (args.log_msm, msm_ms, msm_ok))
    step_ms = msm_ms + ntt_ms
    n_ntt = 1 << args.log_ntt
    ntt_gbs = 64.0 * n_ntt / (ntt_ms * 1e-3) / 1e9
    out = {
        "metric": "ms per BN254 MSM/NTT at 2^24; batched jobs/s + p99 vs one-at-a-time calls",
        "value": round(step_ms, 3),
        "unit": "ms per step (one 2^24 G1 MSM + one 2^24 Fr NTT)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 limbs (exact Fq/Fr Montgomery arithmetic)",
        "data": "synthetic, seeded",
        "config": {"workload": "BN254 G1 MSM 2^%d + Fr NTT 2^%d (fwd n2b)" % (args.log_msm, args.log_ntt), "msm_log_n": args.log_msm,
                   "ntt_log_n": args.log_ntt},
        "msm_ms": round(msm_ms, 3),
        "ntt_ms": round(ntt_ms, 3),
        "lde_4_cosets_ms": round(lde_ms, 3),
        "parity": {"msm_closed_form": msm_ok, "ntt_round_trip": ntt_ok},
        "roofline": {"bound": "hbm", "kernel": "ntt_pass_kernel (whole NTT call)", "achieved": round(ntt_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ntt_gbs / HBM_PEAK_GBS, 4), "traffic": None},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(O)
    dev.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
