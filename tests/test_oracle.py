"""Pins the CPU oracle against the reference's golden vectors and against
mathematics that fixes each function independently (Python big integers,
closed forms, brute force).  No GPU needed."""
import json
import os
import random

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
P = 21888242871839275222246405745257275088696311157297823662689037894645226208583
R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
M = 1 << 256


def brev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2) if bits else 0


# ---------------- field ----------------
def test_field_constants_match_python(oracle):
    O = oracle
    one = O.to_limbs([M % R])
    # omega has exact order 2^28 and equals 7^((r-1)/2^28)
    w = O.from_mont(O.root_of_unity(28))[0]
    assert w == pow(7, (R - 1) >> 28, R)
    assert pow(w, 1 << 27, R) == R - 1
    assert O.from_mont(O.root_of_unity(1))[0] == R - 1
    assert O.from_mont(O.root_of_unity(0))[0] == 1
    winv = O.from_mont(O.root_of_unity(28, inverse=True))[0]
    assert winv * w % R == 1
    # Montgomery one round-trips
    assert O.from_mont(one)[0] == 1


@pytest.mark.parametrize("op", ["ax", "a_plus_x", "x_plus_y", "ax_plus_y", "x_minus_y", "ax_minus_y", "x_minus_ay", "x_mul_y"])
def test_ff_ops_vs_bigint(oracle, op):
    O = oracle
    rng = random.Random(op)
    corners = [0, 1, 2, R - 1, R - 2, M % R, (R - 1) // 2, 1 << 253]
    xs = corners + [rng.randrange(R) for _ in range(200)]
    ys = list(reversed(corners)) + [rng.randrange(R) for _ in range(200)]
    a = rng.randrange(R)
    out = O.ff_op(op, a=O.to_mont([a]), x=O.to_mont(xs), y=O.to_mont(ys))
    got = O.from_mont(out)
    f = {
        "ax": lambda x, y: a * x,
        "a_plus_x": lambda x, y: a + x,
        "x_plus_y": lambda x, y: x + y,
        "ax_plus_y": lambda x, y: a * x + y,
        "x_minus_y": lambda x, y: x - y,
        "ax_minus_y": lambda x, y: a * x - y,
        "x_minus_ay": lambda x, y: x - a * y,
        "x_mul_y": lambda x, y: x * y,
    }[op]
    assert got == [f(x, y) % R for x, y in zip(xs, ys)]
    # outputs are canonical limbs
    assert all(v < R for v in O.from_limbs(out))


def test_fq_mul_vs_bigint(oracle):
    O = oracle
    rng = random.Random(5)
    xs = [0, 1, P - 1] + [rng.randrange(P) for _ in range(300)]
    ys = [P - 1, 1, P - 1] + [rng.randrange(P) for _ in range(300)]
    out = O.mul_vec(O.to_mont(xs, P), O.to_mont(ys, P), field="fq")
    assert O.from_mont(out, P) == [x * y % P for x, y in zip(xs, ys)]


def test_batch_inverse_vs_bigint(oracle):
    O = oracle
    rng = random.Random(9)
    xs = [1, 2, R - 1] + [rng.randrange(1, R) for _ in range(100)]
    out = O.from_mont(O.batch_inverse(O.to_mont(xs)))
    assert out == [pow(x, -1, R) for x in xs]
    assert O.from_mont(O.batch_inverse(O.to_mont([0])))[0] == 0


# ---------------- curve ----------------
def py_g1_mul(k):
    """Affine [k]G over Fq with Python integers (independent of the oracle)."""
    def add(p1, p2):
        if p1 is None:
            return p2
        if p2 is None:
            return p1
        (x1, y1), (x2, y2) = p1, p2
        if x1 == x2 and (y1 + y2) % P == 0:
            return None
        if p1 == p2:
            lam = 3 * x1 * x1 * pow(2 * y1, -1, P) % P
        else:
            lam = (y2 - y1) * pow(x2 - x1, -1, P) % P
        x3 = (lam * lam - x1 - x2) % P
        return (x3, (lam * (x1 - x3) - y1) % P)

    acc, base = None, (1, 2)
    while k:
        if k & 1:
            acc = add(acc, base)
        base = add(base, base)
        k >>= 1
    return acc


def test_msm_bases_fixture_matches_oracle_and_python(oracle):
    O = oracle
    data = json.load(open(os.path.join(GOLDEN, "msm_bases.json")))
    pts = [(int(x, 16), int(y, 16)) for x, y in data["points"]]
    assert len(pts) == 256
    for k in (0, 1, 2, 3, 17, 100, 255):
        assert pts[k] == py_g1_mul(k + 1)
        a = O.mul_generator(k + 1)
        assert (O.from_mont(a[:8], P)[0], O.from_mont(a[8:], P)[0]) == pts[k]


def test_curve_is_on_curve_and_group_order(oracle):
    O = oracle
    # [r]G = infinity, [r-1]G = -G
    a = O.mul_generator(R - 1)
    assert (O.from_mont(a[:8], P)[0], O.from_mont(a[8:], P)[0]) == (1, P - 2)


# ---------------- NTT ----------------
@pytest.mark.parametrize("log_n", [1, 2, 3, 5, 8])
@pytest.mark.parametrize("log_ext,coset", [(0, 0), (1, 1), (2, 3), (3, 5)])
def test_ntt_vs_bruteforce_dft(oracle, log_n, log_ext, coset):
    O = oracle
    x = O.random_fr(log_n * 100 + coset, 1 << log_n)
    got = O.ntt(x, log_n, False, False, log_ext, coset)
    # independent big-int evaluation: out[brev(k)] = P(g * w_N^coset * w_n^k)
    coeffs = O.from_mont(x)
    n = 1 << log_n
    w_n = pow(pow(7, (R - 1) >> 28, R), 1 << (28 - log_n), R)
    rho = 1
    if log_ext:
        w_N = pow(pow(7, (R - 1) >> 28, R), 1 << (28 - log_n - log_ext), R)
        rho = 7 * pow(w_N, coset, R) % R
    vals = O.from_mont(got)
    for k in range(n) if n <= 32 else random.Random(1).sample(range(n), 32):
        z = rho * pow(w_n, k, R) % R
        assert vals[brev(k, log_n)] == sum(c * pow(z, i, R) for i, c in enumerate(coeffs)) % R


@pytest.mark.parametrize("log_n", [1, 4, 10])
@pytest.mark.parametrize("bitrev_in", [False, True])
@pytest.mark.parametrize("log_ext,coset", [(0, 0), (2, 1)])
def test_ntt_round_trip_and_orders(oracle, log_n, bitrev_in, log_ext, coset):
    O = oracle
    x = O.random_fr(77 + log_n, 1 << log_n)
    y = O.ntt(x, log_n, bitrev_in, False, log_ext, coset)
    z = O.ntt(y, log_n, not bitrev_in, True, log_ext, coset)
    assert (z == x).all()
    # bit-reversed input variant equals the natural one up to permutations
    if bitrev_in:
        ref = O.ntt(O.bit_reverse(x, log_n), log_n, False, False, log_ext, coset)
        assert (O.bit_reverse(y, log_n) == ref).all()


def test_ntt_lde_known_answers(oracle):
    O = oracle
    kats = json.load(open(os.path.join(GOLDEN, "ntt_lde_kat.json")))
    # lde_2: coset c block c*n..(c+1)*n, bit-reversed inside each block
    k2 = kats["correctness_lde_2"]
    L, e = k2["log_values_count"], k2["log_extension_degree"]
    inp = O.to_limbs([int(v, 16) for v in k2["inputs"]])
    exp = [int(v, 16) for v in k2["expected_outputs"]]
    for c in range(1 << e):
        out = O.from_limbs(O.ntt(inp, L, False, False, e, c))
        assert out == exp[c << L:(c + 1) << L]
    # lde_4: the fixture lists cosets in bit-reversed coset order (whole-domain
    # bit reversal); the kernel's coset c is fixture block brev_2(c).  The
    # reference test asserts coset 0 only (tests/ntt_test.cu:184-185).
    k4 = kats["correctness_lde_4"]
    L, e = k4["log_values_count"], k4["log_extension_degree"]
    inp = O.to_limbs([int(v, 16) for v in k4["inputs"]])
    exp = [int(v, 16) for v in k4["expected_outputs"]]
    for c in range(1 << e):
        out = O.from_limbs(O.ntt(inp, L, False, False, e, c))
        b = brev(c, e)
        assert out == exp[b << L:(b + 1) << L]


# ---------------- MSM ----------------
def _closed_form(O, s_ints, k_mont):
    ks = O.from_mont(k_mont)
    t = sum(a * b for a, b in zip(ks, s_ints)) % R
    return O.mul_generator(t)


@pytest.mark.parametrize("log_n", [0, 2, 5, 8])
def test_msm_pippenger_vs_naive_and_closed_form(oracle, log_n):
    O = oracle
    n = 1 << log_n
    s = O.from_limbs(O.random_fr(1000 + log_n, n, montgomery=False))
    bases = O.gen_bases(s)
    k = O.msm_scalars(2000 + log_n, n)
    res = O.msm(bases, k, log_n)
    folded = O.to_affine(O.msm_fold(res))[0]
    assert (folded == _closed_form(O, s, k)).all()
    if n <= 32:
        assert (O.to_affine(O.msm_naive(bases, k))[0] == folded).all()


@pytest.mark.parametrize("c", [3, 4, 7, 9, 12, 13])
def test_msm_any_window_width_same_fold(oracle, c):
    O = oracle
    n = 64
    s = O.from_limbs(O.random_fr(3000 + c, n, montgomery=False))
    bases = O.gen_bases(s)
    k = O.msm_scalars(4000 + c, n)
    folded = O.to_affine(O.msm_fold(O.msm(bases, k, 6, c)))[0]
    assert (folded == _closed_form(O, s, k)).all()


def test_msm_per_bit_results_definition(oracle):
    """result[w*c+j] = sum of bucket sums whose digit has bit j (j < c-1);
    checked against an independent Python restatement of the digits."""
    O = oracle
    n, c = 16, 5
    s = O.from_limbs(O.random_fr(77, n, montgomery=False))
    bases = O.gen_bases(s)
    k = O.msm_scalars(78, n)
    res = O.msm(bases, k, 4, c)
    W = (254 - 1) // c + 1
    top_signed = c - (W * c - 254) - 1
    coeff = [0] * 254
    for i, kv in enumerate(O.from_mont(k)):
        sign = 0
        if kv >= R - kv and kv != 0:
            kv = R - kv
            sign = 1
        borrow = 0
        for w in range(W):
            d = ((kv >> (w * c)) & ((1 << c) - 1)) + borrow
            sg = sign
            borrow = 0
            if d > (1 << (c - 1)):
                d = (1 << c) - d
                borrow = 1
                sg ^= 1
            if d == 0:
                continue
            sv = -s[i] if sg else s[i]
            top = (1 << top_signed) if w == W - 1 else (1 << (c - 1))
            if d == top:
                coeff[w * c + (top_signed if w == W - 1 else c - 1)] += sv
            else:
                for j in range(c - 1):
                    if (d >> j) & 1:
                        coeff[w * c + j] += sv
    aff = O.to_affine(res)
    for i in range(254):
        t = coeff[i] % R
        if t == 0:
            assert (aff[i] == 0).all()
        else:
            assert (aff[i] == O.mul_generator(t)).all(), i


def test_msm_digits_reconstruct_scalar(oracle):
    O = oracle
    for c in (3, 13, 16, 20, 22):
        k = O.msm_scalars(c, 50)
        W = (254 - 1) // c + 1
        for i, kv in enumerate(O.from_mont(k)):
            d, sg = O.msm_digits(k[i], c)
            v = sum((-int(dd) if s else int(dd)) << (w * c) for w, (dd, s) in enumerate(zip(d, sg)))
            assert v % R == kv
            assert all(dd <= (1 << (c - 1)) for dd in d)


def test_window_bits_table(oracle):
    O = oracle
    table = {0: 3, 1: 3, 3: 3, 10: 10, 12: 12, 13: 13, 14: 13, 15: 15, 19: 15, 20: 16, 21: 16, 22: 17, 23: 19, 24: 20, 25: 20, 26: 22}
    for L, c in table.items():
        assert O.msm_window_bits(L) == c


# ---------------- PN ----------------
def test_pn_closed_form_from_reference_test(oracle):
    """tests/pn_test.cu:128-163: idx(row, col) = (row*cols + col) % 3."""
    O = oracle
    cols, log_rows, m = 4, 10, 3
    rows = 1 << log_rows
    cells = cols * rows
    idx = np.zeros(cells, dtype=np.uint32)
    for row in range(rows):
        for col in range(cols):
            idx[col * rows + row] = (row * cols + col) % m
    ks = [random.Random(1).randrange(R) for _ in range(cols)]
    out = O.from_mont(O.pn(idx, O.to_mont(ks), cols, log_rows))
    w = pow(pow(7, (R - 1) >> 28, R), 1 << (28 - log_rows), R)
    for row in range(rows):
        for col in range(cols):
            s = row * cols + col
            t = s if s % m == 0 else (s + (cells - s - 1) // m * m if s < m else s - m)
            assert out[(t % cols) * rows + t // cols] == pow(w, row, R) * ks[col] % R


def test_pn_random_cycles_is_a_permutation_rotation(oracle):
    O = oracle
    cols, log_rows = 3, 6
    rows = 1 << log_rows
    cells = cols * rows
    rng = random.Random(4)
    perm = list(range(cells))
    rng.shuffle(perm)
    idx = np.zeros(cells, dtype=np.uint32)
    pos, var = 0, 0
    while pos < cells:
        L = rng.randint(1, 9)
        for p in perm[pos:pos + L]:
            idx[p] = var + 5
        pos += L
        var += 1
    ks = [rng.randrange(R) for _ in range(cols)]
    out = O.from_mont(O.pn(idx, O.to_mont(ks), cols, log_rows))
    w = pow(pow(7, (R - 1) >> 28, R), 1 << (28 - log_rows), R)
    ident = {}
    groups = {}
    for row in range(rows):
        for col in range(cols):
            cell = col * rows + row
            ident[cell] = pow(w, row, R) * ks[col] % R
            groups.setdefault(int(idx[cell]), []).append((row * cols + col, cell))
    first = min(groups)
    for key, members in groups.items():
        members.sort()
        cl = [c for _, c in members]
        for j, cell in enumerate(cl):
            nxt = cell if key == first else cl[(j + 1) % len(cl)]
            assert out[cell] == ident[nxt]


def _pn_reference_semantics(idx, ks, cols, log_rows):
    """Pure-Python restatement of src/pn.cu:17-97 for small inputs: a stable
    sort of the row-major cells by the low ceil(log2 cols) + log_rows bits of
    their variable index, run-length encoding by the full index, each run a
    cycle (a cell takes the next cell's identity value), the first run fixed."""
    rows = 1 << log_rows
    bits = (cols - 1).bit_length() + log_rows
    w = pow(pow(7, (R - 1) >> 28, R), 1 << (28 - log_rows), R)
    cells = [(row * cols + col, col * rows + row) for row in range(rows) for col in range(cols)]
    order = sorted(cells, key=lambda t: (int(idx[t[1]]) & ((1 << bits) - 1), t[0]))
    out = [None] * (cols * rows)
    start = 0
    run = 0
    while start < len(order):
        end = start + 1
        while end < len(order) and idx[order[end][1]] == idx[order[start][1]]:
            end += 1
        members = [c for _, c in order[start:end]]
        for j, cell in enumerate(members):
            src = cell if run == 0 else members[(j + 1) % len(members)]
            out[cell] = pow(w, src % rows, R) * ks[src // rows] % R
        start = end
        run += 1
    return out


@pytest.mark.parametrize("cols,log_rows", [(3, 3), (4, 4)])
def test_pn_indexes_beyond_the_sorted_bits(oracle, cols, log_rows):
    """Variable indexes of 2^bits and more (bits = ceil(log2 cols) + log_rows)
    are sorted by their low bits only (src/pn.cu:43-50) while runs compare full
    indexes (src/pn.cu:58-66): a wide index and a narrow one with the same low
    bits interleave and split each other's cells into several runs; a wide
    first index splits the first run.  The oracle follows these rules as the
    independent restatement above does."""
    O = oracle
    rows = 1 << log_rows
    bits = (cols - 1).bit_length() + log_rows
    rng = random.Random(0xBC0000D1 + cols)
    for trial in range(20):
        low = [rng.randrange(4) for _ in range(cols * rows)]
        idx = np.array([v + (rng.randrange(3) << bits) for v in low], dtype=np.uint32)
        if trial % 2:
            idx[0] = (1 << bits) + low[0]  # a wide index on the first row-major cell: the first run is wide
        ks = [rng.randrange(R) for _ in range(cols)]
        got = O.from_mont(O.pn(idx, O.to_mont(ks), cols, log_rows))
        assert got == _pn_reference_semantics(idx, ks, cols, log_rows), trial


# ---------------- ff vector API ----------------
def test_ff_vector_api_closed_forms(oracle):
    O = oracle
    rng = random.Random(11)
    xs = [rng.randrange(1, R) for _ in range(64)]
    gp = O.from_mont(O.grand_product(O.to_mont(xs)))
    acc = 1
    for i, x in enumerate(xs):
        acc = acc * x % R
        assert gp[i] == acc
    b = rng.randrange(R)
    mp = O.from_mont(O.multiply_by_powers(O.to_mont(xs), O.to_mont([b])))
    assert mp == [x * pow(b, i, R) % R for i, x in enumerate(xs)]
    pe = O.from_mont(O.poly_evaluate(O.to_mont(xs), O.to_mont([b])))[0]
    assert pe == sum(x * pow(b, i, R) for i, x in enumerate(xs)) % R
    w10 = pow(pow(7, (R - 1) >> 28, R), 1 << 18, R)
    pw = O.from_mont(O.powers_of_w(10, 5, 40))
    assert pw == [pow(w10, 5 + i, R) for i in range(40)]
    pwi = O.from_mont(O.powers_of_w(10, 1020, 8, inverse=True, bit_reversed=True))
    assert pwi == [pow(w10, -brev((1020 + i) % 1024, 10), R) for i in range(8)]
    pg = O.from_mont(O.powers_of_g(12, 3, 10))
    assert pg == [pow(7, 3 + i, R) for i in range(10)]
    pgi = O.from_mont(O.powers_of_g(12, 3, 10, inverse=True, bit_reversed=True))
    assert pgi == [pow(7, -brev(3 + i, 12), R) for i in range(10)]
    sh = O.from_mont(O.omega_shift(O.to_mont(xs), 10, 3, 7))
    assert sh == [x * pow(w10, 3 * (7 + i), R) % R for i, x in enumerate(xs)]
    br = O.bit_reverse(O.to_mont(xs), 6)
    assert O.from_mont(br) == [xs[brev(i, 6)] for i in range(64)]


def test_test_helpers_vs_bigint(oracle):
    """The oracle helpers the GPU parity tests use as checkers: multi-point
    evaluation (against a Python Horner loop), the Fr dot product of the MSM
    closed form and the progression s0 + i d (against Python integers)."""
    R = oracle.R
    x = oracle.random_fr(0xBC0000C1, 77)
    z = oracle.random_fr(0xBC0000C2, 5)
    xs, zs = oracle.from_mont(x), oracle.from_mont(z)
    got = oracle.from_mont(oracle.poly_eval_many(x, z))
    for zi, g in zip(zs, got):
        acc = 0
        for c in reversed(xs):
            acc = (acc * zi + c) % R
        assert g == acc
    y = oracle.random_fr(0xBC0000C3, 77)
    assert oracle.from_mont(oracle.fr_dot(x, y))[0] == sum(a * b for a, b in zip(xs, oracle.from_mont(y))) % R
    s0, d = R - 5, 0xFEDCBA9876543210FEDCBA
    assert oracle.from_mont(oracle.fr_progression(s0, d, 300)) == [(s0 + i * d) % R for i in range(300)]
