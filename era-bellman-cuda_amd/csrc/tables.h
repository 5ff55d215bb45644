// Device-resident power tables shared by the ff, ntt and pn paths.
//
// A power table answers v^e for e < 2^28 with at most one multiply:
//   v^e = fine[e & (2^14 - 1)] * coarse[e >> 14],
// the same two-level idea as the reference's powers_data (src/ff_kernels.cuh:8-27),
// with a fixed 14/14 split.  Tables exist for omega (order 2^28), the
// generator g = 7 and g^-1 (src/ff_kernels.cu:65-91).
#pragma once

#include "device.cuh"

namespace bc {

constexpr unsigned POW_FINE_BITS = 14;
constexpr unsigned POW_LOG_RANGE = 28;

struct power_table {
  const fe *fine;   // 2^14 entries: v^j
  const fe *coarse; // 2^14 entries: v^(j << 14)
};

struct ff_state {
  power_table w;  // omega_{2^28}
  power_table g;  // 7
  power_table gi; // 7^-1
  fe *storage;    // single allocation backing all tables
  bool ready;
};

// Tables of the current device; nullptr before ff_set_up.
const ff_state *ff_current();

// v^e for e < 2^28 (canonical result)
__device__ __forceinline__ fe table_pow(const power_table &t, unsigned e) {
  const fe c = load_fe(t.coarse + ((e >> POW_FINE_BITS) & ((1u << (POW_LOG_RANGE - POW_FINE_BITS)) - 1)));
  const unsigned f = e & ((1u << POW_FINE_BITS) - 1);
  if (f == 0)
    return c;
  return mul<FrParams>(load_fe(t.fine + f), c);
}

// omega_{2^28}^(-e) = omega^(2^28 - e)
__device__ __forceinline__ unsigned inverse_exponent(unsigned e) { return ((1u << POW_LOG_RANGE) - e) & ((1u << POW_LOG_RANGE) - 1); }

} // namespace bc
