"""CPU ORACLE — test infrastructure only.

ctypes wrapper over liboracle.so (the C restatement of era-bellman-cuda's
host arithmetic; see bn254.h).  Only tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py may import this module, and only as the checker.

Field elements are numpy uint32 arrays of shape (n, 8) (little-endian limbs);
affine G1 points (n, 16) = x||y; Jacobian points (n, 24) = x||y||z.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

P = 21888242871839275222246405745257275088696311157297823662689037894645226208583
R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
MONT = 1 << 256

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("liboracle.so is not built (run make -C oracle)")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.or_splitmix64.restype = ctypes.c_uint64
        _lib.or_msm_window_bits.restype = ctypes.c_uint
        _lib.or_pn.restype = ctypes.c_int
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _fe(n):
    return np.zeros((n, 8), dtype=np.uint32)


# ---- conversions between Python ints and limb arrays ----
def to_limbs(values):
    out = np.zeros((len(values), 8), dtype=np.uint32)
    for i, v in enumerate(values):
        for j in range(8):
            out[i, j] = (v >> (32 * j)) & 0xFFFFFFFF
    return out


def from_limbs(arr):
    arr = np.asarray(arr, dtype=np.uint32).reshape(-1, 8)
    return [sum(int(arr[i, j]) << (32 * j) for j in range(8)) for i in range(arr.shape[0])]


def to_mont(values, m=R):
    return to_limbs([(v * MONT) % m for v in values])


def from_mont(arr, m=R):
    inv = pow(MONT, -1, m)
    return [(v * inv) % m for v in from_limbs(arr)]


# ---- field ----
OP = {"ax": 0, "a_plus_x": 1, "x_plus_y": 2, "ax_plus_y": 3, "x_minus_y": 4, "ax_minus_y": 5, "x_minus_ay": 6, "x_mul_y": 7}


def ff_op(op, a=None, x=None, y=None):
    n = (x if x is not None else y).shape[0]
    out = _fe(n)
    a_ = np.ascontiguousarray(a if a is not None else _fe(1), dtype=np.uint32)
    x_ = np.ascontiguousarray(x if x is not None else _fe(n), dtype=np.uint32)
    y_ = np.ascontiguousarray(y if y is not None else _fe(n), dtype=np.uint32)
    lib().or_ff_op(OP[op], _p(a_), _p(x_), _p(y_), _p(out), ctypes.c_size_t(n))
    return out


def mul_vec(x, y, field="fr"):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    y = np.ascontiguousarray(y, dtype=np.uint32)
    out = _fe(x.shape[0])
    lib().or_mul_vec(1 if field == "fq" else 0, _p(x), _p(y), _p(out), ctypes.c_size_t(x.shape[0]))
    return out


def random_fr(seed, n, montgomery=True):
    out = _fe(n)
    lib().or_random_fr(ctypes.c_uint64(seed), _p(out), ctypes.c_size_t(n), 1 if montgomery else 0)
    return out


def msm_scalars(seed, n):
    out = _fe(n)
    lib().or_msm_scalars(ctypes.c_uint64(seed), _p(out), ctypes.c_size_t(n))
    return out


def set_threads(n):
    lib().or_set_threads(int(n))


def get_threads():
    return int(lib().or_get_threads())


def root_of_unity(log_order, inverse=False):
    out = _fe(1)
    lib().or_root_of_unity(_p(out), log_order, 1 if inverse else 0)
    return out


# ---- NTT ----
def ntt(x, log_n, bit_reversed_inputs=False, inverse=False, log_ext=0, coset=0):
    y = np.ascontiguousarray(x, dtype=np.uint32).copy()
    lib().or_ntt(_p(y), log_n, int(bit_reversed_inputs), int(inverse), log_ext, coset)
    return y


def dft_n2b(x, log_n, log_ext=0, coset=0):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    out = _fe(1 << log_n)
    lib().or_dft_n2b(_p(x), _p(out), log_n, log_ext, coset)
    return out


def poly_eval(coeffs, z):
    c = np.ascontiguousarray(coeffs, dtype=np.uint32)
    zz = np.ascontiguousarray(z, dtype=np.uint32).reshape(1, 8)
    out = _fe(1)
    lib().or_poly_eval(_p(c), c.shape[0], _p(zz), _p(out))
    return out


def poly_eval_many(coeffs, zs):
    """P(z) for each Montgomery point of zs (m, 8), evaluated in parallel."""
    c = np.ascontiguousarray(coeffs, dtype=np.uint32)
    zz = np.ascontiguousarray(zs, dtype=np.uint32).reshape(-1, 8)
    out = _fe(zz.shape[0])
    lib().or_poly_eval_many(_p(c), c.shape[0], _p(zz), zz.shape[0], _p(out))
    return out


# ---- MSM ----
def msm_window_bits(log_n):
    return int(lib().or_msm_window_bits(log_n))


def msm_digits(scalar_mont, c):
    w = (254 - 1) // c + 1
    d = np.zeros(256, dtype=np.uint32)
    s = np.zeros(256, dtype=np.uint32)
    sc = np.ascontiguousarray(scalar_mont, dtype=np.uint32).reshape(1, 8)
    lib().or_msm_digits(_p(sc), c, _p(d), _p(s))
    return d[:w].copy(), s[:w].copy()


def msm(bases, scalars, log_n, c=None):
    if c is None:
        c = msm_window_bits(log_n)
    b = np.ascontiguousarray(bases, dtype=np.uint32)
    s = np.ascontiguousarray(scalars, dtype=np.uint32)
    out = np.zeros((254, 24), dtype=np.uint32)
    lib().or_msm(_p(b), _p(s), log_n, c, _p(out))
    return out


def msm_fold(results):
    r = np.ascontiguousarray(results, dtype=np.uint32).reshape(254, 24)
    out = np.zeros((1, 24), dtype=np.uint32)
    lib().or_msm_fold(_p(r), _p(out))
    return out


def msm_naive(bases, scalars):
    b = np.ascontiguousarray(bases, dtype=np.uint32)
    s = np.ascontiguousarray(scalars, dtype=np.uint32)
    out = np.zeros((1, 24), dtype=np.uint32)
    lib().or_msm_naive(_p(b), _p(s), b.shape[0], _p(out))
    return out


def to_affine(jac):
    j = np.ascontiguousarray(jac, dtype=np.uint32).reshape(-1, 24)
    out = np.zeros((j.shape[0], 16), dtype=np.uint32)
    lib().or_batch_to_affine(_p(j), _p(out), ctypes.c_size_t(j.shape[0]))
    return out


def mul_generator(k):
    """[k]G as canonical affine Montgomery limbs (16,) ; k a Python int."""
    s = to_limbs([k % R])
    jac = np.zeros((1, 24), dtype=np.uint32)
    lib().g1_mul_generator(_p(jac), _p(s))
    return to_affine(jac)[0]


def gen_bases(s_canonical_ints):
    s = to_limbs(s_canonical_ints)
    out = np.zeros((len(s_canonical_ints), 16), dtype=np.uint32)
    lib().or_gen_bases(_p(s), _p(out), len(s_canonical_ints))
    return out


def gen_bases_progression(s0, d, n):
    a = to_limbs([s0 % R])
    b = to_limbs([d % R])
    out = np.zeros((n, 16), dtype=np.uint32)
    lib().or_gen_bases_progression(_p(a), _p(b), _p(out), n)
    return out


# ---- PN ----
def pn(indexes, scalars, cols, log_rows):
    idx = np.ascontiguousarray(indexes, dtype=np.uint32)
    sc = np.ascontiguousarray(scalars, dtype=np.uint32)
    out = _fe(cols << log_rows)
    rc = lib().or_pn(_p(idx), _p(sc), _p(out), cols, log_rows)
    if rc:
        raise ValueError("columns_count must be 3 or 4")
    return out


# ---- ff vector API ----
def grand_product(x):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    out = _fe(x.shape[0])
    lib().or_grand_product(_p(x), _p(out), ctypes.c_size_t(x.shape[0]))
    return out


def multiply_by_powers(x, base):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    b = np.ascontiguousarray(base, dtype=np.uint32).reshape(1, 8)
    out = _fe(x.shape[0])
    lib().or_multiply_by_powers(_p(x), _p(b), _p(out), ctypes.c_size_t(x.shape[0]))
    return out


def batch_inverse(x):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    out = _fe(x.shape[0])
    lib().or_batch_inverse(_p(x), _p(out), ctypes.c_size_t(x.shape[0]))
    return out


def poly_evaluate(x, point):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    pt = np.ascontiguousarray(point, dtype=np.uint32).reshape(1, 8)
    out = _fe(1)
    lib().or_poly_evaluate(_p(x), _p(pt), ctypes.c_size_t(x.shape[0]), _p(out))
    return out


def fr_dot(x, y):
    """sum x_i y_i mod r (Montgomery in, Montgomery out)."""
    x = np.ascontiguousarray(x, dtype=np.uint32)
    y = np.ascontiguousarray(y, dtype=np.uint32)
    assert x.shape == y.shape
    out = _fe(1)
    lib().or_fr_dot(_p(x), _p(y), ctypes.c_size_t(x.shape[0]), _p(out))
    return out


def fr_progression(s0, d, n):
    """(s0 + i d) mod r for i < n, Montgomery limbs; s0, d Python ints."""
    a, b = to_mont([s0 % R]), to_mont([d % R])
    out = _fe(n)
    lib().or_fr_progression(_p(a), _p(b), _p(out), ctypes.c_size_t(n))
    return out


def powers_of_w(log_degree, offset, count, inverse=False, bit_reversed=False):
    out = _fe(count)
    lib().or_powers_of_w(_p(out), log_degree, offset, count, int(inverse), int(bit_reversed))
    return out


def powers_of_g(log_degree, offset, count, inverse=False, bit_reversed=False):
    out = _fe(count)
    lib().or_powers_of_g(_p(out), log_degree, offset, count, int(inverse), int(bit_reversed))
    return out


def omega_shift(values, log_degree, shift, offset, inverse=False):
    v = np.ascontiguousarray(values, dtype=np.uint32)
    out = _fe(v.shape[0])
    lib().or_omega_shift(_p(v), _p(out), log_degree, shift, offset, v.shape[0], int(inverse))
    return out


def bit_reverse(x, log_n):
    x = np.ascontiguousarray(x, dtype=np.uint32)
    out = _fe(1 << log_n)
    lib().or_bit_reverse(_p(x), _p(out), log_n)
    return out
