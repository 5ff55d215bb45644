// Host build of the 29-bit-limb field and curve functions, for the CPU tests
// that check them against Python integers (tests/test_field29_host.py).
#include "../../era-bellman-cuda_amd/csrc/curve29.cuh"

using namespace bc;

namespace {
template <class F> void binary(int op, const uint32_t *a, const uint32_t *b, uint32_t *out) {
  f29 x, y, r;
  for (int i = 0; i < 9; i++) {
    x.v[i] = a[i];
    y.v[i] = b[i];
  }
  switch (op) {
  case 0: r = mul29<F>(x, y); break;
  case 1: r = sqr29<F>(x); break;
  case 2: r = norm29(x); break;
  case 3: r = reduce_value29<F>(x); break;
  case 4: r = add29(x, y); break;
  case 5: r = sub29(x, y, P29<F>::c2); break;
  case 6: r = sub29(x, y, P29<F>::c16); break;
  case 7: r = sub29(x, y, P29<F>::c48e2); break;
  case 8: r = neg29(y, P29<F>::c16); break;
  default: r = x; break;
  }
  for (int i = 0; i < 9; i++)
    out[i] = r.v[i];
}
} // namespace

extern "C" {

// op: 0 mul, 1 sqr, 2 norm, 3 reduce_value, 4 add, 5/6/7 sub with 2m/16m/48m, 8 16m - b
void h29_op(int field, int op, const uint32_t *a, const uint32_t *b, uint32_t *out) {
  if (field)
    binary<FqParams>(op, a, b, out);
  else
    binary<FrParams>(op, a, b, out);
}

void h29_mul2(int field, const uint32_t *a, const uint32_t *b, const uint32_t *c, const uint32_t *d, uint32_t *out) {
  f29 x[4];
  const uint32_t *in[4] = {a, b, c, d};
  for (int k = 0; k < 4; k++)
    for (int i = 0; i < 9; i++)
      x[k].v[i] = in[k][i];
  const f29 r = field ? mul2_29<FqParams>(x[0], x[1], x[2], x[3]) : mul2_29<FrParams>(x[0], x[1], x[2], x[3]);
  for (int i = 0; i < 9; i++)
    out[i] = r.v[i];
}

int h29_is_zero(int field, const uint32_t *a) {
  f29 x;
  for (int i = 0; i < 9; i++)
    x.v[i] = a[i];
  return field ? is_zero29<FqParams>(x) : is_zero29<FrParams>(x);
}

void h29_ext_to29(const uint32_t *a, uint32_t *out) {
  fe x;
  for (int i = 0; i < 8; i++)
    x.v[i] = a[i];
  const f29 r = ext_to29(x);
  for (int i = 0; i < 9; i++)
    out[i] = r.v[i];
}

void h29_to_ext(int field, const uint32_t *a, uint32_t *out) {
  f29 x;
  for (int i = 0; i < 9; i++)
    x.v[i] = a[i];
  const fe r = field ? to_ext<FqParams>(x) : to_ext<FrParams>(x);
  for (int i = 0; i < 8; i++)
    out[i] = r.v[i];
}

void h29_constants(int field, uint32_t *out) {
  // n, one, c2, c16, c48e2, c64 (9 limbs each), inv, pos_inv
  auto put = [&](int k, const c29::limbs &l) {
    for (int i = 0; i < 9; i++)
      out[9 * k + i] = l.v[i];
  };
  if (field) {
    using P = P29<FqParams>;
    put(0, P::n); put(1, P::one); put(2, P::c2); put(3, P::c16); put(4, P::c48e2); put(5, P::c64);
    out[54] = P::inv; out[55] = P::pos_inv;
  } else {
    using P = P29<FrParams>;
    put(0, P::n); put(1, P::one); put(2, P::c2); put(3, P::c16); put(4, P::c48e2); put(5, P::c64);
    out[54] = P::inv; out[55] = P::pos_inv;
  }
}

// sum of `count` XYZZ points (4 x 8 words each, 2^256 radix, below 2p;
// ZZ = 0 is infinity) through xyzz29_accumulate, returned in the 2^256 radix
void h29_add_chain(const uint32_t *points, int count, uint32_t *out, int *out_empty) {
  xyzz29 acc{};
  bool empty = true;
  for (int k = 0; k < count; k++) {
    g1_xyzz q;
    fe *c[4] = {&q.x, &q.y, &q.zz, &q.zzz};
    for (int j = 0; j < 4; j++)
      for (int i = 0; i < 8; i++)
        c[j]->v[i] = points[32 * k + 8 * j + i];
    xyzz29_accumulate(acc, empty, q);
  }
  *out_empty = empty;
  const g1_xyzz r = empty ? xyzz_infinity() : xyzz29_to_xyzz(acc);
  const fe *c[4] = {&r.x, &r.y, &r.zz, &r.zzz};
  for (int k = 0; k < 4; k++)
    for (int i = 0; i < 8; i++)
      out[8 * k + i] = c[k]->v[i];
}

// sum of `count` affine points (2 x 8 words each, canonical 2^256 radix),
// each negated when its flag is set, through xyzz29_add_affine, returned in
// the 2^256 radix.  *out_bounds = 1 when every intermediate sum kept the
// bounds the addition relies on: normalized limbs, X below 2m and Y, ZZ, ZZZ
// below 2^257.3 (checked as below 2^257.3 < 1.2312 * 2^257).
void h29_madd_chain(const uint32_t *points, const int *negate, int count, uint32_t *out, int *out_empty, int *out_bounds) {
  using P = P29<FqParams>;
  xyzz29 acc{};
  bool empty = true;
  bool ok = true;
  auto below = [](const f29 &v, const c29::limbs &bound) { // v < bound, both normalized
    for (int i = 8; i >= 0; i--)
      if (v.v[i] != bound.v[i])
        return v.v[i] < bound.v[i];
    return false;
  };
  auto normalized = [](const f29 &v) {
    for (int i = 0; i < 8; i++)
      if (v.v[i] > MASK29)
        return false;
    return true;
  };
  constexpr c29::limbs two_m = c29::times(P::n, 2);
  c29::limbs big{}; // 1.2312 * 2^257 > 2^257.3
  big.v[8] = (uint32_t)(1.2312 * (double)(1u << 25));
  for (int k = 0; k < count; k++) {
    g1_affine q;
    for (int i = 0; i < 8; i++) {
      q.x.v[i] = points[16 * k + i];
      q.y.v[i] = points[16 * k + 8 + i];
    }
    xyzz29_add_affine(acc, empty, q, negate[k] != 0, &q);
    if (!empty) {
      ok = ok && normalized(acc.x) && normalized(acc.y) && normalized(acc.zz) && normalized(acc.zzz);
      ok = ok && below(acc.x, two_m) && below(acc.y, big) && below(acc.zz, big) && below(acc.zzz, big);
    }
  }
  *out_empty = empty;
  *out_bounds = ok;
  const g1_xyzz r = empty ? xyzz_infinity() : xyzz29_to_xyzz(acc);
  const fe *c[4] = {&r.x, &r.y, &r.zz, &r.zzz};
  for (int k = 0; k < 4; k++)
    for (int i = 0; i < 8; i++)
      out[8 * k + i] = c[k]->v[i];
}

} // extern "C"
