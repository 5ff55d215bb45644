// Fr vector API (ff_*): power tables, element-wise ops, scans, batch
// inverse, powers, shifts, permutations.  Semantics follow src/ff.cu and
// src/ff_kernels.cu; every result is a canonical Montgomery Fr element.
#include <cstring>
#include <mutex>

#include "common.h"
#include "scan.h"
#include "sort.h"
#include "tables.h"

namespace bc {

namespace {

ff_state g_ff[MAX_DEVICES];
std::mutex g_ff_mutex;

// table entries: fine[j] = v^j, coarse[j] = v^(j << 14); one thread per entry
__global__ void build_power_table_kernel(fe *fine, fe *coarse, fe v) {
  const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned n = 1u << POW_FINE_BITS;
  if (j >= 2 * n)
    return;
  const bool is_coarse = j >= n;
  const unsigned e = j & (n - 1);
  // v^(e << 14) = (v^(2^14))^e
  fe base = v;
  if (is_coarse)
    for (unsigned i = 0; i < POW_FINE_BITS; i++)
      base = mul<FrParams>(base, base);
  const fe r = pow_u64<FrParams>(base, e);
  store_fe(is_coarse ? coarse + e : fine + e, r);
}

hipError_t build_table(power_table &t, fe *storage, fe v) {
  fe *fine = storage, *coarse = storage + (1u << POW_FINE_BITS);
  build_power_table_kernel<<<grid_for(2u << POW_FINE_BITS, 256), 256>>>(fine, coarse, v);
  BC_LAUNCH_CHECK();
  t.fine = fine;
  t.coarse = coarse;
  return hipSuccess;
}

} // namespace

const ff_state *ff_current() {
  const int d = device_slot();
  return d >= 0 && g_ff[d].ready ? &g_ff[d] : nullptr;
}

hipError_t ff_set_up_impl() {
  std::lock_guard<std::mutex> lock(g_ff_mutex);
  const int d = device_slot();
  if (d < 0)
    return hipErrorInvalidDevice;
  ff_state &s = g_ff[d];
  if (s.ready)
    return hipSuccess;
  const size_t per_table = 2u << POW_FINE_BITS;
  BC_TRY(hipMalloc(&s.storage, sizeof(fe) * per_table * 3));
  const fe w = fe_const<FrParams>(FrParams::omega);
  fe seven = fe_zero();
  seven.v[0] = FrParams::generator;
  const fe g = to_mont<FrParams>(seven);
  const fe gi = inverse<FrParams>(g);
  BC_TRY(build_table(s.w, s.storage, w));
  BC_TRY(build_table(s.g, s.storage + per_table, g));
  BC_TRY(build_table(s.gi, s.storage + 2 * per_table, gi));
  BC_TRY(hipDeviceSynchronize());
  s.ready = true;
  return hipSuccess;
}

hipError_t ff_tear_down_impl() {
  std::lock_guard<std::mutex> lock(g_ff_mutex);
  const int d = device_slot();
  if (d < 0)
    return hipErrorInvalidDevice;
  ff_state &s = g_ff[d];
  if (!s.ready)
    return hipSuccess;
  BC_TRY(hipDeviceSynchronize());
  BC_TRY(hipFree(s.storage));
  s = ff_state{};
  return hipSuccess;
}

// ---------------------------------------------------------------- element-wise
enum class ew_op { ax, a_plus_x, x_plus_y, ax_plus_y, x_minus_y, ax_minus_y, x_minus_ay, x_mul_y, set };

template <ew_op OP> __device__ __forceinline__ fe ew_apply(const fe &a, const fe *x, const fe *y, size_t i) {
  using F = FrParams;
  switch (OP) {
  case ew_op::ax:
    return mul<F>(a, load_fe(x + i));
  case ew_op::a_plus_x:
    return add<F>(a, load_fe(x + i));
  case ew_op::x_plus_y:
    return add<F>(load_fe(x + i), load_fe(y + i));
  case ew_op::ax_plus_y:
    return add<F>(mul<F>(a, load_fe(x + i)), load_fe(y + i));
  case ew_op::x_minus_y:
    return sub<F>(load_fe(x + i), load_fe(y + i));
  case ew_op::ax_minus_y:
    return sub<F>(mul<F>(a, load_fe(x + i)), load_fe(y + i));
  case ew_op::x_minus_ay:
    return sub<F>(load_fe(x + i), mul<F>(a, load_fe(y + i)));
  case ew_op::x_mul_y:
    return mul<F>(load_fe(x + i), load_fe(y + i));
  case ew_op::set:
  default:
    return a;
  }
}

// `a` is either a kernel argument (host scalar) or read once per thread from device memory.
template <ew_op OP>
__global__ __launch_bounds__(256) void ew_kernel(fe a_val, const fe *a_ref, const fe *__restrict__ x, const fe *__restrict__ y, fe *__restrict__ r,
                                                 unsigned count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count)
    return;
  const fe a = a_ref ? load_fe(a_ref) : a_val;
  store_fe(r + i, ew_apply<OP>(a, x, y, i));
}

template <ew_op OP> hipError_t ew_launch(const void *a, const void *x, const void *y, void *r, unsigned count, hipStream_t stream) {
  if (count == 0)
    return hipSuccess;
  fe a_val = fe_zero();
  const fe *a_ref = nullptr;
  if (a) {
    if (classify_pointer(a) == mem_kind::device)
      a_ref = static_cast<const fe *>(a);
    else
      memcpy(&a_val, a, sizeof(fe));
  }
  ew_kernel<OP><<<grid_for(count, 256), 256, 0, stream>>>(a_val, a_ref, static_cast<const fe *>(x), static_cast<const fe *>(y), static_cast<fe *>(r), count);
  return hipGetLastError();
}

// ---------------------------------------------------------------- powers and shifts
// target[i] = omega_{2^log_degree}^(+-j), j = offset+i (bit-reversed over log_degree bits if asked)
__global__ void powers_of_w_kernel(power_table w, fe *target, unsigned log_degree, unsigned offset, unsigned count, bool inverse, bool bit_reversed) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count)
    return;
  const unsigned raw = offset + i;
  const unsigned idx = bit_reversed ? brev_bits(raw, log_degree) : raw;
  unsigned e = (idx << (POW_LOG_RANGE - log_degree)) & ((1u << POW_LOG_RANGE) - 1);
  if (inverse)
    e = inverse_exponent(e);
  store_fe(target + i, table_pow(w, e));
}

// target[i] = g^(+-j), j = (offset+i or its bit reverse) mod 2^28 — the reference's table range (src/ff_kernels.cu:573-585)
__global__ void powers_of_g_kernel(power_table g, fe *target, unsigned log_degree, unsigned offset, unsigned count, bool bit_reversed) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count)
    return;
  const unsigned raw = offset + i;
  const unsigned idx = bit_reversed ? brev_bits(raw, log_degree) : raw;
  store_fe(target + i, table_pow(g, idx & ((1u << POW_LOG_RANGE) - 1)));
}

// result[i] = values[i] * omega_{2^log_degree}^(+-shift*(offset+i))  (src/ff_kernels.cu:601-621)
__global__ void omega_shift_kernel(power_table w, const fe *values, fe *result, unsigned log_degree, unsigned shift, unsigned offset, unsigned count,
                                   bool inverse) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count)
    return;
  unsigned e = ((shift * (offset + i)) << (POW_LOG_RANGE - log_degree)) & ((1u << POW_LOG_RANGE) - 1);
  if (inverse)
    e = inverse_exponent(e);
  store_fe(result + i, mul<FrParams>(load_fe(values + i), table_pow(w, e)));
}

// ---------------------------------------------------------------- permutations
__global__ void bit_reverse_kernel(const fe *__restrict__ in, fe *__restrict__ out, unsigned log_n) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (1u << log_n))
    return;
  store_fe(out + brev_bits(i, log_n), load_fe(in + i));
}

__global__ void bit_reverse_in_place_kernel(fe *data, unsigned log_n) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (1u << log_n))
    return;
  const unsigned j = brev_bits(i, log_n);
  if (i >= j)
    return;
  const fe a = load_fe(data + i), b = load_fe(data + j);
  store_fe(data + i, b);
  store_fe(data + j, a);
}

__global__ void select_kernel(const fe *__restrict__ src, fe *__restrict__ dst, const unsigned *__restrict__ idx, unsigned count) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count)
    return;
  store_fe(dst + i, load_fe(src + idx[i]));
}

// ---------------------------------------------------------------- batch inverse
// Each thread inverts a contiguous chunk with one field inversion (Montgomery's
// trick); prefix products go to `scratch` (which may alias `out`, not `in`).
constexpr unsigned INV_CHUNK = 32;
__global__ __launch_bounds__(256) void batch_inverse_kernel(const fe *in, fe *scratch, fe *out, unsigned count) {
  using F = FrParams;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t begin = t * INV_CHUNK;
  if (begin >= count)
    return;
  const size_t end = begin + INV_CHUNK < count ? begin + INV_CHUNK : count;
  fe acc = fe_one<F>();
  for (size_t i = begin; i < end; i++) {
    store_fe(scratch + i, acc);
    const fe v = load_fe(in + i);
    if (!fe_is_zero(v))
      acc = mul<F>(acc, v);
  }
  fe inv = inverse<F>(acc);
  for (size_t i = end; i-- > begin;) {
    const fe v = load_fe(in + i);
    const fe pre = load_fe(scratch + i);
    if (fe_is_zero(v)) {
      store_fe(out + i, v);
      continue;
    }
    store_fe(out + i, mul<F>(pre, inv));
    inv = mul<F>(inv, v);
  }
}

// ---------------------------------------------------------------- packed bits / distribute (src/pn_kernels.cu:123-160)
__global__ void set_values_from_packed_bits_kernel(fe *values, const unsigned *bits, unsigned count) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count)
    return;
  const unsigned b = (bits[i >> 5] >> (i & 31)) & 1u;
  store_fe(values + i, b ? fe_one<FrParams>() : fe_zero());
}

__global__ void distribute_values_kernel(const fe *src, fe *dst, unsigned count, unsigned stride) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count)
    return;
  store_fe(dst + (size_t)i * stride, load_fe(src + i));
}

hipError_t copy_scalar_to_device(fe *dst, const void *src, hipStream_t stream) {
  return hipMemcpyAsync(dst, src, sizeof(fe), hipMemcpyDefault, stream);
}

} // namespace bc

using namespace bc;

extern "C" {

bc_error ff_set_up(unsigned powers_of_w_coarse_log_count, unsigned powers_of_g_coarse_log_count) {
  (void)powers_of_w_coarse_log_count;
  (void)powers_of_g_coarse_log_count;
  return as_bc(ff_set_up_impl());
}

bc_error ff_tear_down() { return as_bc(ff_tear_down_impl()); }

bc_error ff_set_value(void *target, const void *value, unsigned count, bc_stream stream) {
  return as_bc(ew_launch<ew_op::set>(value, nullptr, nullptr, target, count, to_stream(stream)));
}

bc_error ff_set_value_zero(void *target, unsigned count, bc_stream stream) {
  if (count == 0)
    return bc_success;
  return as_bc(hipMemsetAsync(target, 0, sizeof(fe) * (size_t)count, to_stream(stream)));
}

bc_error ff_set_value_one(void *target, unsigned count, bc_stream stream) {
  const fe one = fe_one<FrParams>();
  return as_bc(ew_launch<ew_op::set>(&one, nullptr, nullptr, target, count, to_stream(stream)));
}

bc_error ff_ax(const void *a, const void *x, void *result, unsigned count, bc_stream stream) {
  return as_bc(ew_launch<ew_op::ax>(a, x, nullptr, result, count, to_stream(stream)));
}

bc_error ff_a_plus_x(const void *a, const void *x, void *result, unsigned count, bc_stream stream) {
  return as_bc(ew_launch<ew_op::a_plus_x>(a, x, nullptr, result, count, to_stream(stream)));
}

bc_error ff_x_plus_y(const void *x, const void *y, void *result, unsigned count, bc_stream stream) {
  return as_bc(ew_launch<ew_op::x_plus_y>(nullptr, x, y, result, count, to_stream(stream)));
}

bc_error ff_ax_plus_y(const void *a, const void *x, const void *y, void *result, unsigned count, bc_stream stream) {
  return as_bc(ew_launch<ew_op::ax_plus_y>(a, x, y, result, count, to_stream(stream)));
}

bc_error ff_x_minus_y(const void *x, const void *y, void *result, unsigned count, bc_stream stream) {
  return as_bc(ew_launch<ew_op::x_minus_y>(nullptr, x, y, result, count, to_stream(stream)));
}

bc_error ff_ax_minus_y(const void *a, const void *x, const void *y, void *result, unsigned count, bc_stream stream) {
  return as_bc(ew_launch<ew_op::ax_minus_y>(a, x, y, result, count, to_stream(stream)));
}

bc_error ff_x_minus_ay(const void *a, const void *x, const void *y, void *result, unsigned count, bc_stream stream) {
  return as_bc(ew_launch<ew_op::x_minus_ay>(a, x, y, result, count, to_stream(stream)));
}

bc_error ff_x_mul_y(const void *x, const void *y, void *result, unsigned count, bc_stream stream) {
  return as_bc(ew_launch<ew_op::x_mul_y>(nullptr, x, y, result, count, to_stream(stream)));
}

// Inclusive prefix product (src/ff.cu:104-118).
bc_error ff_grand_product(ff_grand_product_configuration cfg) {
  arena ar = arena::pool_mode(to_pool(cfg.mem_pool), to_stream(cfg.stream));
  return as_bc(prefix_product(static_cast<const fe *>(cfg.inputs), static_cast<fe *>(cfg.outputs), cfg.count, ar));
}

// outputs[i] = inputs[i] * base^i via a prefix product of [1, base, base, ...] (src/ff.cu:136-163).
bc_error ff_multiply_by_powers(ff_multiply_by_powers_configuration cfg) {
  hipStream_t s = to_stream(cfg.stream);
  hipMemPool_t pool = to_pool(cfg.mem_pool);
  const unsigned n = cfg.count;
  if (n == 0)
    return bc_success;
  fe *powers = nullptr;
  const bool use_outputs = cfg.inputs != cfg.outputs;
  if (use_outputs)
    powers = static_cast<fe *>(cfg.outputs);
  else
    BC_TRY_BC(pool_alloc(&powers, n, pool, s));
  BC_TRY_BC(ff_set_value_one(powers, 1, cfg.stream));
  if (n > 1) {
    BC_TRY_BC(ff_set_value(powers + 1, cfg.base, n - 1, cfg.stream));
    arena ar = arena::pool_mode(pool, s);
    BC_TRY_BC(prefix_product(powers + 1, powers + 1, n - 1, ar));
  }
  BC_TRY_BC(ff_x_mul_y(cfg.inputs, powers, cfg.outputs, n, cfg.stream));
  if (!use_outputs)
    BC_TRY_BC(pool_free(powers, s));
  return bc_success;
}

bc_error ff_inverse(ff_inverse_configuration cfg) {
  hipStream_t s = to_stream(cfg.stream);
  const unsigned n = cfg.count;
  if (n == 0)
    return bc_success;
  fe *in = static_cast<fe *>(cfg.inputs), *out = static_cast<fe *>(cfg.outputs);
  fe *scratch = out;
  if (in == out)
    BC_TRY_BC(pool_alloc(&scratch, n, to_pool(cfg.mem_pool), s));
  const unsigned threads = (n + INV_CHUNK - 1) / INV_CHUNK;
  batch_inverse_kernel<<<grid_for(threads, 256), 256, 0, s>>>(in, scratch, out, n);
  BC_TRY_BC(hipGetLastError());
  if (in == out)
    BC_TRY_BC(pool_free(scratch, s));
  return bc_success;
}

// sum_i values[i] * point^i (src/ff.cu:183-201); point and result may be host or device pointers.
bc_error ff_poly_evaluate(ff_poly_evaluate_configuration cfg) {
  hipStream_t s = to_stream(cfg.stream);
  hipMemPool_t pool = to_pool(cfg.mem_pool);
  const unsigned n = cfg.count;
  fe *temp = nullptr, *sum = nullptr;
  BC_TRY_BC(pool_alloc(&temp, n ? n : 1, pool, s));
  BC_TRY_BC(pool_alloc(&sum, 1, pool, s));
  if (n) {
    ff_multiply_by_powers_configuration mp{cfg.mem_pool, cfg.stream, cfg.values, cfg.point, temp, n};
    BC_TRY_BC(ff_multiply_by_powers(mp));
  }
  arena ar = arena::pool_mode(pool, s);
  BC_TRY_BC(field_sum(temp, sum, n, ar));
  BC_TRY_BC(hipMemcpyAsync(cfg.result, sum, sizeof(fe), hipMemcpyDefault, s));
  BC_TRY_BC(pool_free(temp, s));
  BC_TRY_BC(pool_free(sum, s));
  return bc_success;
}

bc_error ff_get_powers_of_w(void *target, unsigned log_degree, unsigned offset, unsigned count, bool inverse, bool bit_reversed, bc_stream stream) {
  const ff_state *st = ff_current();
  if (!st || log_degree > POW_LOG_RANGE)
    return bc_error_invalid_value;
  if (count == 0)
    return bc_success;
  powers_of_w_kernel<<<grid_for(count, 256), 256, 0, to_stream(stream)>>>(st->w, static_cast<fe *>(target), log_degree, offset, count, inverse, bit_reversed);
  return as_bc(hipGetLastError());
}

bc_error ff_get_powers_of_g(void *target, unsigned log_degree, unsigned offset, unsigned count, bool inverse, bool bit_reversed, bc_stream stream) {
  const ff_state *st = ff_current();
  if (!st || log_degree > 32)
    return bc_error_invalid_value;
  if (count == 0)
    return bc_success;
  powers_of_g_kernel<<<grid_for(count, 256), 256, 0, to_stream(stream)>>>(inverse ? st->gi : st->g, static_cast<fe *>(target), log_degree, offset, count,
                                                                          bit_reversed);
  return as_bc(hipGetLastError());
}

bc_error ff_omega_shift(const void *values, void *result, unsigned log_degree, unsigned shift, unsigned offset, unsigned count, bool inverse,
                        bc_stream stream) {
  const ff_state *st = ff_current();
  if (!st || log_degree > POW_LOG_RANGE)
    return bc_error_invalid_value;
  if (count == 0)
    return bc_success;
  omega_shift_kernel<<<grid_for(count, 256), 256, 0, to_stream(stream)>>>(st->w, static_cast<const fe *>(values), static_cast<fe *>(result), log_degree,
                                                                          shift, offset, count, inverse);
  return as_bc(hipGetLastError());
}

bc_error ff_bit_reverse(const void *values, void *result, unsigned log_count, bc_stream stream) {
  if (log_count > 31)
    return bc_error_invalid_value;
  const unsigned n = 1u << log_count;
  hipStream_t s = to_stream(stream);
  if (values == result)
    bit_reverse_in_place_kernel<<<grid_for(n, 256), 256, 0, s>>>(static_cast<fe *>(result), log_count);
  else
    bit_reverse_kernel<<<grid_for(n, 256), 256, 0, s>>>(static_cast<const fe *>(values), static_cast<fe *>(result), log_count);
  return as_bc(hipGetLastError());
}

bc_error ff_bit_reverse_multigpu(const void **values, void **results, unsigned log_count, const bc_stream *streams, const int *device_ids,
                                 unsigned log_devices_count) {
  if (log_devices_count != 0)
    return bc_error_invalid_value;
  int prev = 0;
  BC_TRY_BC(hipGetDevice(&prev));
  BC_TRY_BC(hipSetDevice(device_ids[0]));
  bc_error e = ff_bit_reverse(values[0], results[0], log_count, streams[0]);
  (void)hipSetDevice(prev);
  return e;
}

bc_error ff_select(const void *source, void *destination, const unsigned *indexes, unsigned count, bc_stream stream) {
  if (count == 0)
    return bc_success;
  select_kernel<<<grid_for(count, 256), 256, 0, to_stream(stream)>>>(static_cast<const fe *>(source), static_cast<fe *>(destination), indexes, count);
  return as_bc(hipGetLastError());
}

bc_error ff_sort_u32(ff_sort_u32_configuration cfg) {
  arena ar = arena::pool_mode(to_pool(cfg.mem_pool), to_stream(cfg.stream));
  return as_bc(radix_sort_keys(static_cast<const uint32_t *>(cfg.values), static_cast<uint32_t *>(cfg.sorted_values), cfg.count, 32, ar));
}

bc_error pn_set_values_from_packed_bits(void *values, const void *packet_bits, unsigned count, bc_stream stream) {
  if (count == 0)
    return bc_success;
  set_values_from_packed_bits_kernel<<<grid_for(count, 256), 256, 0, to_stream(stream)>>>(static_cast<fe *>(values), static_cast<const unsigned *>(packet_bits),
                                                                                          count);
  return as_bc(hipGetLastError());
}

bc_error pn_distribute_values(const void *src, void *dst, unsigned count, unsigned stride, bc_stream stream) {
  if (count == 0)
    return bc_success;
  distribute_values_kernel<<<grid_for(count, 256), 256, 0, to_stream(stream)>>>(static_cast<const fe *>(src), static_cast<fe *>(dst), count, stride);
  return as_bc(hipGetLastError());
}

} // extern "C"
