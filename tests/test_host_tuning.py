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


def test_msm_tuning_builtin_table(B):
    """A fresh process starts from the built-in table: measured (window,
    entries per accumulating thread) for 2^12..2^24, the plain call's window
    width (src/msm.cu:12-37) above."""
    import subprocess
    import sys

    code = ("import ctypes, bellman_amd as B\n"
            "for L in range(0, 29):\n"
            "    t = B.bc_amd_msm_tuning(); B.call('bc_amd_msm_get_tuning', ctypes.c_uint(L), ctypes.byref(t)); print(L, t.window_bits, t.piece)\n")
    pkg = os.path.dirname(os.path.dirname(B.__file__))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True, env=dict(os.environ, PYTHONPATH=pkg))
    table = {int(a): (int(b), int(c)) for a, b, c in (line.split() for line in out.stdout.splitlines())}
    assert table[12] == (9, 32) and table[14] == (11, 16) and table[16] == (11, 8) and table[20] == (15, 16) and table[24] == (17, 128)
    assert table[10] == (5, 16) and table[9] == (4, 16) and table[0] == (3, 16)
    assert table[25] == (20, 128) and table[26] == (22, 128) and table[28] == (22, 128)
    for L, (c, piece) in table.items():
        assert 3 <= c <= 22 and piece in (8, 16, 32, 64, 128)
        assert (254 - 1) // c + 1 <= 85  # windows of at least 3 bits


def test_tuning_validation(B):
    lib = B.lib()
    assert lib.bc_amd_msm_set_tuning(ctypes.c_uint(29), B.bc_amd_msm_tuning(10, 32)) == B.bc_error_invalid_value
    assert lib.bc_amd_msm_set_tuning(ctypes.c_uint(10), B.bc_amd_msm_tuning(2, 32)) == B.bc_error_invalid_value
    assert lib.bc_amd_msm_set_tuning(ctypes.c_uint(10), B.bc_amd_msm_tuning(25, 32)) == B.bc_error_invalid_value
    assert lib.bc_amd_msm_set_tuning(ctypes.c_uint(10), B.bc_amd_msm_tuning(10, 0)) == B.bc_error_invalid_value
    assert lib.bc_amd_msm_set_tuning(ctypes.c_uint(10), B.bc_amd_msm_tuning(10, 1 << 16)) == B.bc_error_invalid_value
    # sizes of two or more passes only; every pass at most 10 stages
    assert lib.bc_amd_ntt_set_split(ctypes.c_uint(10), ctypes.c_uint(5)) == B.bc_error_invalid_value  # one pass
    assert lib.bc_amd_ntt_set_split(ctypes.c_uint(24), ctypes.c_uint(11)) == B.bc_error_invalid_value
    assert lib.bc_amd_ntt_set_split(ctypes.c_uint(24), ctypes.c_uint(3)) == B.bc_error_invalid_value  # 21 left for 2 passes
    assert lib.bc_amd_ntt_set_split(ctypes.c_uint(20), ctypes.c_uint(9)) == B.bc_error_invalid_value
    assert lib.bc_amd_ntt_set_split(ctypes.c_uint(20), ctypes.c_uint(0)) == B.bc_success
    assert lib.bc_amd_ntt_set_split(ctypes.c_uint(24), ctypes.c_uint(4)) == B.bc_success
    assert lib.bc_amd_ntt_set_split(ctypes.c_uint(24), ctypes.c_uint(0)) == B.bc_success
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
