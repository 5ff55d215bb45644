"""bc_amd_msm_execute_no_alloc: the MSM entry point that allocates nothing
(its temporaries live in a fixed per-GPU device array).  It runs the plain
call's device path, so its 254 per-bit results must equal msm_execute_async's
byte for byte.  Two calls in a row, the second smaller, check that the array
keeps no state a later call depends on; the folds are also checked against
the closed form [sum k_i s_i]G of progression bases."""
import ctypes

import numpy as np
import pytest

from gpu_helpers import msm

pytestmark = pytest.mark.gpu

MAX_LOG = 17  # BC_AMD_MSM_NO_ALLOC_MAX_LOG in include/bellman-amd.h


def no_alloc(B, dev, d_bases, d_scalars, d_results, log_n):
    cfg = B.msm_configuration()
    cfg.stream = dev.stream
    cfg.bases, cfg.scalars, cfg.results = d_bases, d_scalars, d_results
    cfg.log_scalars_count = log_n
    return B.lib().bc_amd_msm_execute_no_alloc(cfg)


def closed_form(oracle, s0, d, n, k_mont):
    ks = oracle.from_mont(k_mont)
    total = sum(kk * ((s0 + i * d) % oracle.R) for i, kk in enumerate(ks)) % oracle.R
    return oracle.mul_generator(total)


@pytest.mark.timeout(300)
def test_no_alloc_matches_plain_call_twice(bellman, oracle):
    B, dev = bellman
    s0, d = 0x5EED0007, 0x1F2E3D4C
    bases = oracle.gen_bases_progression(s0, d, 1 << MAX_LOG)
    db = dev.upload(bases)
    d_r = dev.malloc(254 * 96)
    try:
        for log_n, seed in ((MAX_LOG, 0xBC0000A1), (11, 0xBC0000A2)):
            n = 1 << log_n
            k = oracle.msm_scalars(seed, n)
            print("no-alloc MSM 2^%d (seed %#x)" % (log_n, seed), flush=True)
            expected = msm(dev, db, k, log_n)
            d_k = dev.upload(k)
            # every result must be written by the call: start from a fill pattern
            B.call("bc_memset_async", ctypes.c_void_p(d_r), ctypes.c_int(0x5A), ctypes.c_size_t(254 * 96), dev.stream)
            assert no_alloc(B, dev, db, d_k, d_r, log_n) == B.bc_success
            dev.sync()
            got = dev.download(d_r, (254, 24))
            dev.free(d_k)
            assert np.array_equal(got, expected), "per-bit results differ from msm_execute_async at 2^%d" % log_n
            assert np.array_equal(oracle.to_affine(oracle.msm_fold(got))[0], closed_form(oracle, s0, d, n, k))
    finally:
        dev.free(db)
        dev.free(d_r)


def test_no_alloc_rejects_oversize_and_host_buffers(bellman, oracle):
    B, dev = bellman
    n = 1 << 4
    k = oracle.msm_scalars(0xBC0000A3, n)
    bases = oracle.gen_bases_progression(3, 5, n)
    db, d_k, d_r = dev.upload(bases), dev.upload(k), dev.malloc(254 * 96)
    try:
        # one past the array's limit: rejected before anything is launched
        assert no_alloc(B, dev, db, d_k, d_r, MAX_LOG + 1) == B.bc_error_invalid_value
        # host scalars or host results would need staging memory
        assert no_alloc(B, dev, db, k.ctypes.data, d_r, 4) == B.bc_error_invalid_value
        host_r = np.zeros((254, 24), dtype=np.uint32)
        assert no_alloc(B, dev, db, d_k, host_r.ctypes.data, 4) == B.bc_error_invalid_value
        assert no_alloc(B, dev, db, d_k, d_r, 4) == B.bc_success
        dev.sync()
        assert np.array_equal(dev.download(d_r, (254, 24)), msm(dev, db, k, 4))
    finally:
        for p in (db, d_k, d_r):
            dev.free(p)
