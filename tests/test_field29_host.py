"""The 29-bit-limb field arithmetic (csrc/field29.cuh) and the bucket addition
built on it (csrc/curve29.cuh), compiled for the host and checked against
Python integers at random and extreme inputs: every function's value, its
result bound and the limb bounds the others rely on.  The same source runs
in the MSM accumulation kernel."""
import ctypes
import os
import random
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "tests", "host29", "harness.hip")
HIPCC = "/opt/rocm/bin/hipcc"

P = 21888242871839275222246405745257275088696311157297823662689037894645226208583
R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
M29 = (1 << 29) - 1
FIELDS = {1: P, 0: R}


@pytest.fixture(scope="module")
def h29(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("h29") / "libh29.so")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950", HARNESS, "-o", out], check=True,
                   capture_output=True)
    lib = ctypes.CDLL(out)
    lib.h29_is_zero.restype = ctypes.c_int
    return lib


def limbs(v, n=9, bits=29):
    return np.array([(v >> (bits * i)) & ((1 << bits) - 1) for i in range(n - 1)] + [v >> (bits * (n - 1))], dtype=np.uint32)


def value(l, bits=29):
    return sum(int(x) << (bits * i) for i, x in enumerate(l))


def words(v):
    return np.array([(v >> (32 * i)) & 0xFFFFFFFF for i in range(8)], dtype=np.uint32)


def wvalue(w):
    return sum(int(x) << (32 * i) for i, x in enumerate(w))


def p_(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def op(lib, field, code, a, b=None):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    b = np.ascontiguousarray(a if b is None else b, dtype=np.uint32)
    out = np.zeros(9, dtype=np.uint32)
    lib.h29_op(field, code, p_(a), p_(b), p_(out))
    return out


def normalized(l):
    return all(int(x) <= M29 for x in l[:8]) and int(l[8]) < (1 << 29)


def samples(rng, hi_bits, count):
    """Values below 2^hi_bits: random ones and the extremes."""
    out = [0, 1, (1 << hi_bits) - 1, (1 << hi_bits) - 2]
    out += [rng.getrandbits(hi_bits) for _ in range(count)]
    return out


def test_constants(h29):
    for field, m in FIELDS.items():
        c = np.zeros(56, dtype=np.uint32)
        h29.h29_constants(field, p_(c))
        n, one, c2, c16, c48, c64 = (c[9 * k:9 * k + 9] for k in range(6))
        assert value(n) == m and normalized(n)
        assert value(one) == (1 << 261) % m
        for k, cc, e in ((2, c2, 0), (16, c16, 0), (48, c48, 2), (64, c64, 1)):
            assert value(cc) == k * m
            assert all(int(x) >= (1 << (29 + e)) and int(x) < (1 << (29 + e)) + (1 << 29) for x in cc[:8])
        # 64m - y for any y below 32m (a canonical coordinate times 2^5) keeps a positive top limb
        assert int(c64[8]) >= (32 * m) >> 232
        assert (int(c[54]) * m) % (1 << 29) == M29  # -m^-1
        assert (int(c[55]) * m) % (1 << 29) == 1


def test_conversions(h29):
    rng = random.Random(0xBC2901)
    for a in samples(rng, 256, 200):
        out = np.zeros(9, dtype=np.uint32)
        h29.h29_ext_to29(p_(words(a)), p_(out))
        assert value(out) == a << 5 and normalized(out)
    for field, m in FIELDS.items():
        inv32 = pow(32, -1, m)
        for v in samples(rng, 258, 200):
            out = np.zeros(8, dtype=np.uint32)
            h29.h29_to_ext(field, p_(limbs(v)), p_(out))
            w = wvalue(out)
            assert w % m == v * inv32 % m and w < 2 * m


@pytest.mark.parametrize("field", [1, 0])
def test_products(h29, field):
    m = FIELDS[field]
    rinv = pow(1 << 261, -1, m)
    rng = random.Random(0xBC2902 + field)
    vals = samples(rng, 259, 120) + [m - 1, m, 2 * m - 1]
    for _ in range(300):
        a, b = rng.choice(vals), rng.choice(vals)
        for code, expect in ((0, a * b), (1, a * a)):
            r = op(h29, field, code, limbs(a), limbs(b))
            assert value(r) % m == expect * rinv % m
            assert normalized(r)
            assert value(r) < (a * (b if code == 0 else a) >> 261) + m
    # one operand with unnormalized limbs up to 2^31 (a sum or difference), the other normalized
    for _ in range(100):
        a_l = np.array([rng.randrange(1 << 31) for _ in range(8)] + [rng.randrange(1 << 27)], dtype=np.uint32)
        b = rng.choice(vals)
        r = op(h29, field, 0, a_l, limbs(b))
        assert value(r) % m == value(a_l) * b * rinv % m and normalized(r)
    worst = np.array([(1 << 31) - 1] * 8 + [(1 << 27) - 1], dtype=np.uint32)
    r = op(h29, field, 0, worst, limbs((1 << 259) - 1))
    assert value(r) % m == value(worst) * ((1 << 259) - 1) * rinv % m
    for _ in range(200):
        a, b, c, d = (rng.choice(vals) for _ in range(4))
        out = np.zeros(9, dtype=np.uint32)
        h29.h29_mul2(field, p_(limbs(a)), p_(limbs(b)), p_(limbs(c)), p_(limbs(d)), p_(out))
        assert value(out) % m == (a * b + c * d) * rinv % m and normalized(out)
        assert value(out) < ((a * b + c * d) >> 261) + m


@pytest.mark.parametrize("field", [1, 0])
def test_limb_operations(h29, field):
    m = FIELDS[field]
    rng = random.Random(0xBC2903 + field)
    for _ in range(300):
        # norm: limbs below 2^32 with a total below 2^262
        l = np.array([rng.randrange(1 << 32) for _ in range(8)] + [rng.randrange(1 << 28)], dtype=np.uint32)
        r = op(h29, field, 2, l)
        assert value(r) == value(l) and all(int(x) <= M29 for x in r[:8])
        # reduce_value: normalized below 2^262
        v = rng.randrange(1 << 262)
        r = op(h29, field, 3, limbs(v))
        assert value(r) % m == v % m and value(r) < m + (m >> 10) and normalized(r)
        # sub with 2m (b < 1.01 m), 16m (b < 2^257.3), 48m (b limbs < 2^31, value < 2^258.9)
        a = rng.randrange(1 << 258)
        for code, k, bmax in ((5, 2, m + (m >> 10)), (6, 16, int(2 ** 257.3)), (8, 16, int(2 ** 257.3))):
            b = rng.randrange(bmax)
            r = op(h29, field, code, limbs(a), limbs(b))
            want = (a if code != 8 else 0) + k * m - b
            assert value(r) == want
            assert all(int(x) < (1 << 32) for x in r)
        b_l = np.array([rng.randrange(3 << 29) for _ in range(8)] + [rng.randrange(1 << 26)], dtype=np.uint32)
        r = op(h29, field, 7, limbs(a), b_l)
        assert value(r) == a + 48 * m - value(b_l) and all(int(x) < (1 << 32) for x in r)
    for j in list(range(0, 339, 7)) + [1, 2, 337]:  # every multiple of m below 2^262
        assert h29.h29_is_zero(field, p_(limbs(j * m))) == 1
    for v in [1, m - 1, m + 1, 2 * m + 5] + [rng.randrange(1 << 258) for _ in range(200)]:
        assert h29.h29_is_zero(field, p_(limbs(v))) == (1 if v % m == 0 else 0)


def _affine_add(a, b):
    if a is None:
        return b
    if b is None:
        return a
    (x1, y1), (x2, y2) = a, b
    if x1 == x2:
        if (y1 + y2) % P == 0:
            return None
        lam = 3 * x1 * x1 * pow(2 * y1, -1, P) % P
    else:
        lam = (y2 - y1) * pow(x2 - x1, -1, P) % P
    x3 = (lam * lam - x1 - x2) % P
    return x3, (lam * (x1 - x3) - y1) % P


def test_xyzz_sums_against_affine_arithmetic(h29, oracle):
    """Chains of general XYZZ additions (the bucket-reduction step): points in
    random XYZZ representations (x = X/ZZ, y = Y/ZZZ with random Z, some
    coordinates given as value + p), infinity terms, a repeated point (the
    doubling path) and a point with its negation: the sum equals the affine
    sum computed with Python integers."""
    rng = random.Random(0xBC2905)
    s = [rng.randrange(1, R) for _ in range(16)]
    pts = oracle.gen_bases(s)
    mont = 1 << 256
    mont_inv = pow(mont, -1, P)

    def decode(pt):
        return wvalue(pt[:8]) * mont_inv % P, wvalue(pt[8:16]) * mont_inv % P

    def xyzz_words(xy, neg=False, lazy=False):
        if xy is None:
            return np.zeros(32, dtype=np.uint32)
        x, y = xy
        if neg:
            y = (P - y) % P
        z = rng.randrange(1, P)
        coords = [x * z * z % P, y * z * z * z % P, z * z % P, z * z * z % P]
        words_ = []
        for k, c in enumerate(coords):
            v = c * mont % P
            if lazy and k % 2 == 0 and v + P < (1 << 256):
                v += P
            words_.append(words(v))
        return np.concatenate(words_)

    chains = [
        [(0, 0), (1, 0), (2, 1), (3, 0), (4, 1), (5, 0)],
        [(6, 0), (6, 0)],                      # doubling
        [(7, 0), (7, 1), (8, 0)],              # infinity mid-chain
        [(9, 0), (None, 0), (10, 1), (None, 0)],  # infinity terms
        [(None, 0), (11, 0), (12, 0), (11, 1), (12, 1)],
        [(13, 0), (14, 0), (15, 1), (13, 0)],
    ]
    for chain in chains:
        flat, want = [], None
        for j, (i, ng) in enumerate(chain):
            xy = None if i is None else decode(pts[i])
            flat.append(xyzz_words(xy, ng, lazy=j % 2 == 1))
            if xy is not None:
                want = _affine_add(want, (xy[0], (P - xy[1]) % P if ng else xy[1]))
        arr = np.ascontiguousarray(np.stack(flat), dtype=np.uint32)
        out = np.zeros(32, dtype=np.uint32)
        empty = ctypes.c_int()
        h29.h29_add_chain(p_(arr), len(chain), p_(out), ctypes.byref(empty))
        if want is None:
            assert empty.value == 1
            continue
        assert empty.value == 0
        X, Y, ZZ, ZZZ = (wvalue(out[8 * k:8 * k + 8]) * mont_inv % P for k in range(4))
        assert ZZ and pow(ZZ, 3, P) == pow(ZZZ, 2, P)
        assert X * pow(ZZ, -1, P) % P == want[0]
        assert Y * pow(ZZZ, -1, P) % P == want[1]


def test_affine_bucket_sums_against_affine_arithmetic(h29, oracle):
    """Chains of mixed additions bucket += (x, +-y) (the accumulation step):
    random points, a repeated point (the doubling path), a point followed by
    its negation (infinity mid-chain, then a fresh start), and a long chain;
    the sum equals the affine sum computed with Python integers, and every
    intermediate sum keeps the bounds the next addition relies on."""
    rng = random.Random(0xBC2906)
    s = [rng.randrange(1, R) for _ in range(24)]
    pts = oracle.gen_bases(s)
    mont = 1 << 256
    mont_inv = pow(mont, -1, P)

    def decode(pt):
        return wvalue(pt[:8]) * mont_inv % P, wvalue(pt[8:16]) * mont_inv % P

    chains = [
        [(0, 0), (1, 0), (2, 1), (3, 0), (4, 1), (5, 0)],
        [(6, 0), (6, 0)],                     # doubling
        [(6, 1), (6, 1), (7, 0)],             # doubling of a negated point
        [(7, 0), (7, 1), (8, 0), (9, 1)],     # infinity mid-chain
        [(10, 0), (11, 0), (10, 1), (11, 1)],  # back to infinity
        [(12, 0), (13, 0), (12, 0), (12, 0)],  # a point met again later
        [(rng.randrange(24), rng.randrange(2)) for _ in range(60)],
    ]
    for chain in chains:
        arr = np.ascontiguousarray(np.stack([np.asarray(pts[i], dtype=np.uint32) for i, _ in chain]), dtype=np.uint32)
        neg = (ctypes.c_int * len(chain))(*[ng for _, ng in chain])
        want = None
        for i, ng in chain:
            x, y = decode(pts[i])
            want = _affine_add(want, (x, (P - y) % P if ng else y))
        out = np.zeros(32, dtype=np.uint32)
        empty, bounds = ctypes.c_int(), ctypes.c_int()
        h29.h29_madd_chain(p_(arr), neg, len(chain), p_(out), ctypes.byref(empty), ctypes.byref(bounds))
        assert bounds.value == 1, chain
        if want is None:
            assert empty.value == 1
            continue
        assert empty.value == 0
        X, Y, ZZ, ZZZ = (wvalue(out[8 * k:8 * k + 8]) * mont_inv % P for k in range(4))
        assert ZZ and pow(ZZ, 3, P) == pow(ZZZ, 2, P)
        assert X * pow(ZZ, -1, P) % P == want[0]
        assert Y * pow(ZZZ, -1, P) % P == want[1]
