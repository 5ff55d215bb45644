"""ff_* vector API and pn_* parity through the C ABI against the CPU oracle
and the reference's closed forms (tests/ff_test.cu, tests/pn_test.cu)."""
import ctypes
import random

import numpy as np
import pytest

import bellman_amd as B
from gpu_helpers import ff_binary, u, vp

pytestmark = pytest.mark.gpu

R = 21888242871839275222246405745257275088548364400416034343698204186575808495617

OPS = {
    "ff_ax": ("ax", True, True, False),
    "ff_a_plus_x": ("a_plus_x", True, True, False),
    "ff_x_plus_y": ("x_plus_y", False, True, True),
    "ff_ax_plus_y": ("ax_plus_y", True, True, True),
    "ff_x_minus_y": ("x_minus_y", False, True, True),
    "ff_ax_minus_y": ("ax_minus_y", True, True, True),
    "ff_x_minus_ay": ("x_minus_ay", True, True, True),
    "ff_x_mul_y": ("x_mul_y", False, True, True),
}


# the scalar `a` by host and by device pointer, for the ops that take one
CASES = [(name, on_dev) for name in sorted(OPS) for on_dev in ((False, True) if OPS[name][1] else (False,))]


@pytest.mark.parametrize("name,a_on_device", CASES)
def test_elementwise_ops(bellman, oracle, name, a_on_device):
    _, dev = bellman
    op, has_a, has_x, has_y = OPS[name]
    n = (1 << 16) + 3
    corners = oracle.to_mont([0, 1, R - 1, R - 2, (R - 1) // 2])
    x = np.concatenate([corners, oracle.random_fr(1, n - 5)])
    y = np.concatenate([corners[::-1], oracle.random_fr(2, n - 5)])
    a = oracle.random_fr(3, 1)
    got = ff_binary(dev, name, a if has_a else None, x, y if has_y else None, a_on_device=a_on_device)
    exp = oracle.ff_op(op, a=a, x=x, y=y if has_y else None)
    assert np.array_equal(got, exp)


def test_set_value_zero_one(bellman, oracle):
    _, dev = bellman
    n = 1000
    d = dev.malloc(n * 32)
    v = oracle.random_fr(9, 1)
    B.call("ff_set_value", vp(d), B.ptr(v), u(n), dev.stream)
    dev.sync()
    assert (dev.download(d, (n, 8)) == v[0]).all()
    dv = dev.upload(oracle.random_fr(10, 1))
    B.call("ff_set_value", vp(d), vp(dv), u(n), dev.stream)
    dev.sync()
    assert (dev.download(d, (n, 8)) == dev.download(dv, (1, 8))[0]).all()
    B.call("ff_set_value_one", vp(d), u(n), dev.stream)
    dev.sync()
    assert oracle.from_mont(dev.download(d, (n, 8))) == [1] * n
    B.call("ff_set_value_zero", vp(d), u(n), dev.stream)
    dev.sync()
    assert (dev.download(d, (n, 8)) == 0).all()
    dev.free(d)
    dev.free(dv)


@pytest.mark.parametrize("n", [1, 5, 1024, 1025, 100000])
def test_grand_product_and_powers(bellman, oracle, n):
    _, dev = bellman
    x = oracle.random_fr(n, n)
    dx = dev.upload(x)
    dy = dev.malloc(n * 32)
    B.call("ff_grand_product", B.ff_grand_product_configuration(dev.pool, dev.stream, dx, dy, n))
    dev.sync()
    assert np.array_equal(dev.download(dy, (n, 8)), oracle.grand_product(x))
    base = oracle.random_fr(n + 1, 1)
    B.call("ff_multiply_by_powers", B.ff_multiply_by_powers_configuration(dev.pool, dev.stream, dx, base.ctypes.data, dy, n))
    dev.sync()
    assert np.array_equal(dev.download(dy, (n, 8)), oracle.multiply_by_powers(x, base))
    res = np.zeros((1, 8), dtype=np.uint32)
    B.call("ff_poly_evaluate", B.ff_poly_evaluate_configuration(dev.pool, dev.stream, dx, base.ctypes.data, res.ctypes.data, n))
    dev.sync()
    assert np.array_equal(res, oracle.poly_evaluate(x, base))
    dev.free(dx)
    dev.free(dy)


@pytest.mark.parametrize("n", [1, 31, 33, 4096 + 7])
@pytest.mark.parametrize("in_place", [False, True])
def test_batch_inverse(bellman, oracle, n, in_place):
    _, dev = bellman
    x = oracle.random_fr(n + 50, n)
    x[0] = 0  # zero maps to zero
    dx = dev.upload(x)
    dy = dx if in_place else dev.malloc(n * 32)
    B.call("ff_inverse", B.ff_inverse_configuration(dev.pool, dev.stream, dx, dy, n))
    dev.sync()
    assert np.array_equal(dev.download(dy, (n, 8)), oracle.batch_inverse(x))
    dev.free(dx)
    if not in_place:
        dev.free(dy)


@pytest.mark.parametrize("log_degree,offset,count", [(10, 0, 1024), (16, 1000, 5000), (28, 7, 300), (0, 0, 4)])
@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("bitrev", [False, True])
def test_powers_of_w_and_g(bellman, oracle, log_degree, offset, count, inverse, bitrev):
    _, dev = bellman
    d = dev.malloc(count * 32)
    B.call("ff_get_powers_of_w", vp(d), u(log_degree), u(offset), u(count), ctypes.c_bool(inverse), ctypes.c_bool(bitrev), dev.stream)
    dev.sync()
    assert np.array_equal(dev.download(d, (count, 8)), oracle.powers_of_w(log_degree, offset, count, inverse, bitrev))
    B.call("ff_get_powers_of_g", vp(d), u(log_degree), u(offset), u(count), ctypes.c_bool(inverse), ctypes.c_bool(bitrev), dev.stream)
    dev.sync()
    assert np.array_equal(dev.download(d, (count, 8)), oracle.powers_of_g(log_degree, offset, count, inverse, bitrev))
    dev.free(d)


@pytest.mark.parametrize("log_degree,shift,offset", [(10, 3, 7), (20, 1, 0), (16, 12345, 99)])
@pytest.mark.parametrize("inverse", [False, True])
def test_omega_shift(bellman, oracle, log_degree, shift, offset, inverse):
    _, dev = bellman
    n = 3000
    x = oracle.random_fr(shift, n)
    dx = dev.upload(x)
    dy = dev.malloc(n * 32)
    B.call("ff_omega_shift", vp(dx), vp(dy), u(log_degree), u(shift), u(offset), u(n), ctypes.c_bool(inverse), dev.stream)
    dev.sync()
    assert np.array_equal(dev.download(dy, (n, 8)), oracle.omega_shift(x, log_degree, shift, offset, inverse))
    dev.free(dx)
    dev.free(dy)


@pytest.mark.parametrize("log_n", [0, 1, 7, 16])
def test_bit_reverse_and_select(bellman, oracle, log_n):
    _, dev = bellman
    n = 1 << log_n
    x = oracle.random_fr(log_n + 3, n)
    dx = dev.upload(x)
    dy = dev.malloc(n * 32)
    B.call("ff_bit_reverse", vp(dx), vp(dy), u(log_n), dev.stream)
    B.call("ff_bit_reverse", vp(dx), vp(dx), u(log_n), dev.stream)
    dev.sync()
    exp = oracle.bit_reverse(x, log_n)
    assert np.array_equal(dev.download(dy, (n, 8)), exp)
    assert np.array_equal(dev.download(dx, (n, 8)), exp)
    idx = np.array([random.Random(1).randrange(n) for _ in range(2 * n)], dtype=np.uint32)
    di = dev.upload(idx)
    dz = dev.malloc(2 * n * 32)
    B.call("ff_select", vp(dy), vp(dz), vp(di), u(2 * n), dev.stream)
    dev.sync()
    assert np.array_equal(dev.download(dz, (2 * n, 8)), exp[idx])
    for p in (dx, dy, di, dz):
        dev.free(p)


@pytest.mark.parametrize("n", [1, 100, 5000, 1 << 18])
def test_sort_u32(bellman, n):
    _, dev = bellman
    rng = np.random.default_rng(n)
    v = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
    v[: n // 3] = v[0]
    dv = dev.upload(v)
    ds = dev.malloc(n * 4)
    B.call("ff_sort_u32", B.ff_sort_u32_configuration(dev.pool, dev.stream, dv, ds, n))
    dev.sync()
    assert np.array_equal(dev.download(ds, (n,)), np.sort(v))
    dev.free(dv)
    dev.free(ds)


def test_packed_bits_and_distribute(bellman, oracle):
    _, dev = bellman
    count = 1000
    bits = np.random.default_rng(3).integers(0, 2**32, size=(count + 31) // 32, dtype=np.uint64).astype(np.uint32)
    db = dev.upload(bits)
    dv = dev.malloc(count * 32)
    B.call("pn_set_values_from_packed_bits", vp(dv), vp(db), u(count), dev.stream)
    dev.sync()
    got = oracle.from_mont(dev.download(dv, (count, 8)))
    assert got == [(int(bits[i // 32]) >> (i % 32)) & 1 for i in range(count)]
    x = oracle.random_fr(4, 100)
    dx = dev.upload(x)
    dd = dev.malloc(100 * 3 * 32)
    B.call("ff_set_value_zero", vp(dd), u(300), dev.stream)
    B.call("pn_distribute_values", vp(dx), vp(dd), u(100), u(3), dev.stream)
    dev.sync()
    out = dev.download(dd, (300, 8))
    assert np.array_equal(out[::3], x)
    assert (out[1::3] == 0).all()
    for p in (db, dv, dx, dd):
        dev.free(p)


def _pn(dev, idx, scalars, cols, log_rows):
    di = dev.upload(idx)
    ds = dev.upload(scalars)
    dt = dev.malloc(idx.shape[0] * 32)
    B.call("pn_generate_permutation_polynomials", B.generate_permutation_polynomials_configuration(dev.pool, dev.stream, di, ds, dt, cols, log_rows))
    dev.sync()
    out = dev.download(dt, (idx.shape[0], 8))
    for p in (di, ds, dt):
        dev.free(p)
    return out


@pytest.mark.parametrize("cols", [3, 4])
def test_pn_closed_form(bellman, oracle, cols):
    """tests/pn_test.cu:128-163 with idx(row, col) = (row*cols + col) % 3."""
    _, dev = bellman
    log_rows = 10
    rows = 1 << log_rows
    idx = np.zeros(cols * rows, dtype=np.uint32)
    for row in range(rows):
        for col in range(cols):
            idx[col * rows + row] = (row * cols + col) % 3
    ks = oracle.random_fr(cols, cols)
    got = _pn(dev, idx, ks, cols, log_rows)
    assert np.array_equal(got, oracle.pn(idx, ks, cols, log_rows))


@pytest.mark.parametrize("cols,log_rows", [(4, 12), (3, 9), (4, 16)])
def test_pn_random_cycles(bellman, oracle, cols, log_rows):
    _, dev = bellman
    cells = cols << log_rows
    rng = random.Random(cells)
    perm = list(range(cells))
    rng.shuffle(perm)
    idx = np.zeros(cells, dtype=np.uint32)
    pos = var = 0
    while pos < cells:
        L = rng.randint(1, 64)
        idx[perm[pos:pos + L]] = var
        pos += L
        var += 1
    ks = oracle.random_fr(cells, cols)
    assert np.array_equal(_pn(dev, idx, ks, cols, log_rows), oracle.pn(idx, ks, cols, log_rows))


@pytest.mark.parametrize("cols,log_rows", [(3, 8), (4, 11)])
def test_pn_sparse_variable_indexes(bellman, oracle, cols, log_rows):
    """Variable indexes with gaps anywhere below 2^(2 + log_rows), the smallest
    not 0: three one-cell cycles, one cycle over a third of the cells, then
    random lengths."""
    _, dev = bellman
    cells = cols << log_rows
    rng = random.Random(7 * cells + 1)
    perm = list(range(cells))
    rng.shuffle(perm)
    keys = rng.sample(range(1, 4 << log_rows), cells // 4)
    idx = np.zeros(cells, dtype=np.uint32)
    pos = v = 0
    while pos < cells:
        L = [1, 1, 1, cells // 3][v] if v < 4 else rng.randint(1, 40)
        if v == len(keys) - 1:
            L = cells - pos
        idx[perm[pos:pos + L]] = keys[v]
        pos += L
        v += 1
    assert idx.min() > 0
    ks = oracle.random_fr(cells + 1, cols)
    assert np.array_equal(_pn(dev, idx, ks, cols, log_rows), oracle.pn(idx, ks, cols, log_rows))


@pytest.mark.parametrize("cols,log_rows,pattern", [(4, 6, "one_variable"), (3, 7, "one_variable_nonzero"), (4, 6, "all_distinct"),
                                                    (3, 5, "all_distinct"), (4, 0, "one_variable"), (3, 0, "random"), (4, 1, "random"),
                                                    (3, 2, "random")])
def test_pn_degenerate_indexes(bellman, oracle, cols, log_rows, pattern):
    """Edge cases of the index matrix: one variable in every cell (a single run,
    the first, which maps to itself), every cell its own variable (runs of one
    cell), and matrices of one, two and four rows.  The first two are the
    identity permutation, so their values are omega^row * k_col whatever the
    variable numbering: both must equal the oracle and each other."""
    _, dev = bellman
    cells = cols << log_rows
    rng = random.Random(31 * cells + log_rows)
    if pattern == "one_variable":
        idx = np.zeros(cells, dtype=np.uint32)
    elif pattern == "one_variable_nonzero":
        idx = np.full(cells, 5, dtype=np.uint32)
    elif pattern == "all_distinct":
        idx = np.array(rng.sample(range(cells), cells), dtype=np.uint32)
    else:
        perm = list(range(cells))
        rng.shuffle(perm)
        idx = np.zeros(cells, dtype=np.uint32)
        pos = var = 0
        while pos < cells:
            L = rng.randint(1, 4)
            idx[perm[pos:pos + L]] = var
            pos += L
            var += 1
    ks = oracle.random_fr(cells + 2, cols)
    got = _pn(dev, idx, ks, cols, log_rows)
    assert np.array_equal(got, oracle.pn(idx, ks, cols, log_rows))
    if pattern != "random":
        other = np.arange(cells, dtype=np.uint32) if pattern != "all_distinct" else np.zeros(cells, dtype=np.uint32)
        assert np.array_equal(got, _pn(dev, other, ks, cols, log_rows))


@pytest.mark.parametrize("cols,log_rows,mode", [(3, 6, "interleaved"), (4, 9, "interleaved"), (4, 8, "wide_first"), (3, 10, "two_keys_one_bucket"),
                                                 (3, 10, "long_wide_run"), (4, 12, "random_wide")])
def test_pn_indexes_beyond_the_sorted_bits(bellman, oracle, cols, log_rows, mode):
    """Variable indexes of 2^bits and more (bits = 2 + log_rows): the sort
    orders them by their low bits only, so a wide index and a narrow one with
    the same low bits interleave and form several runs each, as in the
    reference (src/pn.cu:43-66).  Runs are found deterministically: the
    output equals the oracle's, twice in a row."""
    _, dev = bellman
    cells = cols << log_rows
    bits = 2 + log_rows
    rng = random.Random(0xBC0000D2 + cells)
    if mode == "two_keys_one_bucket":
        idx = np.full(cells, (5 << bits) + 3, dtype=np.uint32)  # a wide variable and a narrow one, same low bits
        idx[cells // 2:] = 3
    elif mode == "long_wide_run":
        # the first cell's narrow index shares its low bits with every other cell's wide
        # index: one split bucket whose second run spans all other cells
        idx = np.full(cells, (1 << bits) + 7, dtype=np.uint32)
        idx[0] = 7
    elif mode == "random_wide":
        idx = np.array([rng.randrange(1 << 31) for _ in range(cells)], dtype=np.uint32)
        for j in rng.sample(range(cells), cells // 8):
            idx[j] = (idx[rng.randrange(cells)] & ((1 << bits) - 1)) + (rng.randrange(4) << bits)
    else:
        low = [rng.randrange(16) for _ in range(cells)]
        idx = np.array([v + (rng.randrange(3) << bits) for v in low], dtype=np.uint32)
        if mode == "wide_first":
            idx[0] = (2 << bits) + low[0]
    ks = oracle.random_fr(cells + 3, cols)
    want = oracle.pn(idx, ks, cols, log_rows)
    assert np.array_equal(_pn(dev, idx, ks, cols, log_rows), want)
    assert np.array_equal(_pn(dev, idx, ks, cols, log_rows), want)


def test_pn_served_job_size(bellman, oracle):
    """The served mix's permutation job at full size: 4 columns x 2^20 rows,
    a seeded shuffle of the 2^22 cells cut into cycles of length 1..64, each
    cycle one variable index."""
    _, dev = bellman
    cols, log_rows = 4, 20
    cells = cols << log_rows
    rng = np.random.default_rng(0xBC000005)
    perm = rng.permutation(cells)
    lens = rng.integers(1, 65, size=cells)
    idx = np.zeros(cells, dtype=np.uint32)
    pos = var = 0
    while pos < cells:
        L = int(lens[var])
        idx[perm[pos:pos + L]] = var
        pos += L
        var += 1
    ks = oracle.random_fr(0xBC000005, cols)
    assert np.array_equal(_pn(dev, idx, ks, cols, log_rows), oracle.pn(idx, ks, cols, log_rows))


def test_pn_rejects_other_column_counts(bellman):
    _, dev = bellman
    cfg = B.generate_permutation_polynomials_configuration(dev.pool, dev.stream, None, None, None, 5, 4)
    assert B.lib().pn_generate_permutation_polynomials(cfg) == B.bc_error_invalid_value
