// Extension entry points (include/bellman-amd.h).
//
// bc_amd_generator_multiples: fixed-base comb over bytes.  The table holds
// T[j][v] = v * 2^(8j) * G for 32 byte positions and v in [1, 256); a scalar
// costs at most 32 mixed additions plus one field inversion for the affine
// output.  The table is built once per device.
//
// bc_amd_msm_fold: the caller-side fold of the 254 per-bit MSM results
// (sum = 2*sum + results[i], src/bellman-cuda.h:496-505) on the device, one
// thread per MSM, with the affine normalisation.
#include <mutex>

#include "common.h"
#include "curve.cuh"
#include "tables.h"

namespace bc {

namespace {

constexpr unsigned COMB_BYTES = 32;
constexpr unsigned COMB_ENTRIES = 256;
g1_affine *g_comb[MAX_DEVICES];
std::mutex g_comb_mutex;

__global__ __launch_bounds__(128) void comb_table_kernel(const g1_affine *base_per_byte, g1_affine *table) {
  const unsigned id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= COMB_BYTES * COMB_ENTRIES)
    return;
  const unsigned j = id / COMB_ENTRIES, v = id % COMB_ENTRIES;
  const g1_affine b = {load_fe(&base_per_byte[j].x), load_fe(&base_per_byte[j].y)};
  g1_xyzz acc = xyzz_infinity();
  for (int bit = 7; bit >= 0; bit--) {
    acc = xyzz_dbl(acc);
    if ((v >> bit) & 1)
      acc = xyzz_add_affine(acc, b);
  }
  const g1_affine a = xyzz_to_affine(acc);
  store_fe(&table[id].x, a.x);
  store_fe(&table[id].y, a.y);
}

__global__ __launch_bounds__(128) void generator_multiples_kernel(const fe *scalars, const g1_affine *table, g1_affine *out, unsigned count) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count)
    return;
  const fe s = load_fe(scalars + i);
  g1_xyzz acc = xyzz_infinity();
  for (unsigned j = 0; j < COMB_BYTES; j++) {
    const unsigned v = (s.v[j >> 2] >> (8 * (j & 3))) & 0xff;
    if (v)
      acc = xyzz_add_affine(acc, load_affine(table + j * COMB_ENTRIES + v));
  }
  const g1_affine a = xyzz_to_affine(acc);
  store_fe(&out[i].x, a.x);
  store_fe(&out[i].y, a.y);
}

constexpr unsigned FOLD_MAX = 64; // MSMs per fold launch

struct fold_args {
  const g1_jacobian *results[FOLD_MAX];
};

// Horner over the result bits in XYZZ coordinates: 253 doublings and 254
// additions per MSM, then one inversion.
__global__ __launch_bounds__(FOLD_MAX) void msm_fold_kernel(const fold_args a, g1_affine *out, unsigned count) {
  const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count)
    return;
  const g1_jacobian *r = a.results[j];
  g1_xyzz acc = xyzz_infinity();
  for (int i = 253; i >= 0; i--) {
    acc = xyzz_dbl(acc);
    const fe z = load_fe(&r[i].z);
    const fe zz = mul_lazy<Fq>(z, z);
    const g1_xyzz t = {load_fe(&r[i].x), load_fe(&r[i].y), zz, mul_lazy<Fq>(zz, z)};
    acc = xyzz_add(acc, t);
  }
  const g1_affine p = xyzz_to_affine(acc);
  store_fe(&out[j].x, p.x);
  store_fe(&out[j].y, p.y);
}

hipError_t comb_table(g1_affine **table) {
  std::lock_guard<std::mutex> lock(g_comb_mutex);
  const int d = device_slot();
  if (d < 0)
    return hipErrorInvalidDevice;
  if (!g_comb[d]) {
    // 2^(8j) G for j < 32, computed on the host
    g1_affine bases[COMB_BYTES];
    fe one = fe_zero(), two = fe_zero();
    one.v[0] = 1;
    two.v[0] = 2;
    g1_xyzz p = xyzz_from_affine({to_mont<Fq>(one), to_mont<Fq>(two)});
    for (unsigned j = 0; j < COMB_BYTES; j++) {
      bases[j] = xyzz_to_affine(p);
      for (int k = 0; k < 8; k++)
        p = xyzz_dbl(p);
    }
    g1_affine *dev_bases = nullptr, *t = nullptr;
    BC_TRY(hipMalloc(&dev_bases, sizeof(bases)));
    BC_TRY(hipMalloc(&t, sizeof(g1_affine) * COMB_BYTES * COMB_ENTRIES));
    BC_TRY(hipMemcpy(dev_bases, bases, sizeof(bases), hipMemcpyHostToDevice));
    comb_table_kernel<<<grid_for(COMB_BYTES * COMB_ENTRIES, 128), 128>>>(dev_bases, t);
    BC_LAUNCH_CHECK();
    BC_TRY(hipDeviceSynchronize());
    BC_TRY(hipFree(dev_bases));
    g_comb[d] = t;
  }
  *table = g_comb[d];
  return hipSuccess;
}

} // namespace

} // namespace bc

using namespace bc;

extern "C" {

bc_error bc_amd_generator_multiples(const void *scalars, void *bases, unsigned count, bc_stream stream) {
  if (count == 0)
    return bc_success;
  g1_affine *table = nullptr;
  BC_TRY_BC(comb_table(&table));
  generator_multiples_kernel<<<grid_for(count, 128), 128, 0, to_stream(stream)>>>(static_cast<const fe *>(scalars), table, static_cast<g1_affine *>(bases),
                                                                                  count);
  return as_bc(hipGetLastError());
}

bc_error bc_amd_msm_fold(const void *const *results, void *points, unsigned count, bc_stream stream) {
  if (count && (!results || !points))
    return bc_error_invalid_value;
  for (unsigned base = 0; base < count; base += FOLD_MAX) {
    const unsigned k = count - base < FOLD_MAX ? count - base : FOLD_MAX;
    fold_args a{};
    for (unsigned j = 0; j < k; j++) {
      if (!results[base + j])
        return bc_error_invalid_value;
      a.results[j] = static_cast<const g1_jacobian *>(results[base + j]);
    }
    msm_fold_kernel<<<1, FOLD_MAX, 0, to_stream(stream)>>>(a, static_cast<g1_affine *>(points) + base, k);
    BC_TRY_BC(hipGetLastError());
  }
  return bc_success;
}

} // extern "C"
