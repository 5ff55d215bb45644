// Fr NTT for gfx950.
//
// Semantics (src/ntt_kernels.cu:119-294, src/bellman-cuda.h:516-541):
//   bit_reversed_inputs = false: natural order in, bit-reversed order out;
//   bit_reversed_inputs = true:  bit-reversed in, natural out;
//   forward coset (log_extension_degree > 0): coefficient i is multiplied by
//   (g * omega_{n*2^e}^coset)^i before the transform, g = 7;
//   inverse: omega^-1 twiddles, then n^-1 and (with a coset) the inverse twist
//   on the natural output index; outputs canonical in [0, r).
//
// Design.  The radix-2 butterfly network of stage s pairs elements
// n/2^(s+1) apart, and block b of stage s uses w = omega_{2^(s+1)}^brev_s(b).
// One twiddle table TW[2^s + b] (n entries) therefore serves every size and
// both orders: natural->bit-reversed runs Cooley-Tukey butterflies over
// s = 0..L-1; bit-reversed->natural runs the transposed network, Gentleman-
// Sande butterflies (E,O) -> (E+O, w(E-O)) over s = L-1..0.  A pass executes a
// run of consecutive stages [s_lo, s_hi) on LDS tiles of up to 4096 elements
// (128 KiB): the stages' butterflies stay inside groups of 2^(s_hi-s_lo)
// elements, strided by 2^(L-s_hi).  log n <= 24 takes at most two HBM passes.
// Tiles are mapped XCD-major so that neighbouring strided columns run on the
// same XCD and share its L2 lines.
#include <mutex>
#include <vector>

#include "common.h"
#include "tables.h"

namespace bc {

namespace {

constexpr unsigned NTT_MAX_LOG = 28;
constexpr unsigned NTT_LOG_TILE = 12; // 4096 elements per tile
constexpr unsigned NTT_BLOCK = 512;
constexpr unsigned TWIST_FINE_BITS = 12;

struct twiddle_state {
  fe *fwd = nullptr; // TW[2^s + b] = omega_{2^(s+1)}^brev_s(b), index 0 unused
  fe *inv = nullptr; // inverse powers, same layout
  unsigned log_size = 0;
  std::vector<fe *> retired;
};

constexpr int MAX_DEVICES = 16;
twiddle_state g_tw[MAX_DEVICES];
std::mutex g_tw_mutex;
bool g_attr_set[MAX_DEVICES];

int device_index() {
  int d = 0;
  (void)hipGetDevice(&d);
  return d < MAX_DEVICES ? d : 0;
}

__global__ void twiddle_build_kernel(power_table w, fe *fwd, fe *inv, unsigned log_size) {
  const unsigned m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= (1u << log_size))
    return;
  if (m == 0) {
    store_fe(fwd, fe_one<FrParams>());
    store_fe(inv, fe_one<FrParams>());
    return;
  }
  const unsigned s = 31 - __clz(m);
  const unsigned b = m - (1u << s);
  // omega_{2^(s+1)} = omega_{2^28}^(2^(27-s))
  const unsigned e = brev_bits(b, s) << (27 - s);
  store_fe(fwd + m, table_pow(w, e));
  store_fe(inv + m, table_pow(w, inverse_exponent(e)));
}

// Grow the twiddle tables to cover 2^log_n; older tables stay alive (in-flight
// work may still read them) until ntt_tear_down.
hipError_t ensure_twiddles(unsigned log_n, hipStream_t stream, const fe **fwd, const fe **inv) {
  std::lock_guard<std::mutex> lock(g_tw_mutex);
  twiddle_state &t = g_tw[device_index()];
  if (t.log_size < log_n) {
    const ff_state *ff = ff_current();
    if (!ff)
      return hipErrorNotInitialized;
    const unsigned log_size = log_n < 12 ? 12 : log_n;
    fe *f = nullptr, *i = nullptr;
    BC_TRY(hipMalloc(&f, sizeof(fe) << log_size));
    BC_TRY(hipMalloc(&i, sizeof(fe) << log_size));
    twiddle_build_kernel<<<grid_for(1u << log_size, 256), 256, 0, stream>>>(ff->w, f, i, log_size);
    BC_LAUNCH_CHECK();
    // other streams may use the tables as soon as they are published
    BC_TRY(hipStreamSynchronize(stream));
    if (t.fwd) {
      t.retired.push_back(t.fwd);
      t.retired.push_back(t.inv);
    }
    t.fwd = f;
    t.inv = i;
    t.log_size = log_size;
  }
  *fwd = t.fwd;
  *inv = t.inv;
  return hipSuccess;
}

enum : unsigned { POST_LAZY = 0, POST_REDUCE = 1, POST_SCALE = 2, POST_TABLE = 3 };

struct pass_args {
  const fe *in;
  fe *out;
  const fe *tw;
  unsigned log_n, s_lo, s_hi;
  unsigned log_t, log_c, log_g; // tile = 2^log_g groups x 2^log_t rows x 2^log_c columns
  unsigned tiles;
  unsigned xcd_remap; // tiles divisible by 8
  bool pre_table;     // multiply element i by twist(natural index)
  unsigned post;      // POST_*
  fe scale;           // POST_SCALE constant
  const fe *twist_fine;
  const fe *twist_coarse;
};

__device__ __forceinline__ fe twist_pow(const pass_args &a, unsigned i) {
  const fe c = load_fe(a.twist_coarse + (i >> TWIST_FINE_BITS));
  const unsigned f = i & ((1u << TWIST_FINE_BITS) - 1);
  return f ? mul<FrParams>(load_fe(a.twist_fine + f), c) : c;
}

template <bool GS>
__global__ __launch_bounds__(NTT_BLOCK) void ntt_pass_kernel(const pass_args a) {
  using F = FrParams;
  extern __shared__ fe tile[];
  unsigned tid_tile = blockIdx.x;
  if (a.xcd_remap)
    tid_tile = (blockIdx.x & 7) * (a.tiles >> 3) + (blockIdx.x >> 3);
  const unsigned L = a.log_n;
  const unsigned log_lo = L - a.s_hi;           // bits below the group
  const unsigned lo_tiles_log = log_lo - a.log_c; // tiles along the low bits
  const unsigned lo0 = (tid_tile & ((1u << lo_tiles_log) - 1)) << a.log_c;
  const unsigned hi0 = (tid_tile >> lo_tiles_log) << a.log_g;
  const unsigned elems = 1u << (a.log_g + a.log_t + a.log_c);
  const unsigned cmask = (1u << a.log_c) - 1, tmask = (1u << a.log_t) - 1;

  for (unsigned e = threadIdx.x; e < elems; e += NTT_BLOCK) {
    const unsigned c = e & cmask, t = (e >> a.log_c) & tmask, g = e >> (a.log_c + a.log_t);
    const unsigned idx = ((hi0 + g) << (L - a.s_lo)) | (t << log_lo) | (lo0 + c);
    fe v = load_fe(a.in + idx);
    if (a.pre_table)
      v = mul_lazy<F>(v, twist_pow(a, GS ? brev_bits(idx, L) : idx));
    tile[e] = v;
  }
  __syncthreads();

  // Stages in network order (CT: s_lo..s_hi-1, GS: s_hi-1..s_lo).  An odd
  // count starts with one radix-2 step; the rest go two at a time: a thread
  // holds the 4 elements t0, t0+h/2, t0+h, t0+3h/2 (h = row distance of the
  // lower stage s) and applies stages s and s+1 in registers.
  const unsigned nst = a.s_hi - a.s_lo;
  unsigned done = 0;
  if (nst & 1) {
    const unsigned s = GS ? a.s_hi - 1 : a.s_lo;
    const unsigned log_h = a.s_hi - s - 1;
    const unsigned h = 1u << log_h;
    for (unsigned q = threadIdx.x; q < (elems >> 1); q += NTT_BLOCK) {
      const unsigned c = q & cmask;
      const unsigned r = (q >> a.log_c) & (tmask >> 1);
      const unsigned g = q >> (a.log_c + a.log_t - 1);
      const unsigned t0 = ((r >> log_h) << (log_h + 1)) | (r & (h - 1));
      const unsigned i0 = (((g << a.log_t) | t0) << a.log_c) | c;
      const unsigned i1 = i0 + (h << a.log_c);
      const unsigned b = ((hi0 + g) << (s - a.s_lo)) | (t0 >> (a.s_hi - s));
      const fe E = tile[i0], O = tile[i1];
      if (!GS) {
        const fe wo = s != 0 ? mul_lazy<F>(O, load_fe(a.tw + (1u << s) + b)) : O;
        tile[i0] = add2<F>(E, wo);
        tile[i1] = sub2<F>(E, wo);
      } else {
        const fe d = sub2<F>(E, O);
        tile[i0] = add2<F>(E, O);
        tile[i1] = s != 0 ? mul_lazy<F>(d, load_fe(a.tw + (1u << s) + b)) : d;
      }
    }
    __syncthreads();
    done = 1;
  }
  for (; done < nst; done += 2) {
    // the lower-numbered stage of the pair
    const unsigned s = GS ? a.s_hi - 2 - done : a.s_lo + done;
    const unsigned log_h = a.s_hi - s - 1; // >= 1
    const unsigned h = 1u << log_h, hh = h >> 1;
    for (unsigned q = threadIdx.x; q < (elems >> 2); q += NTT_BLOCK) {
      const unsigned c = q & cmask;
      const unsigned u = (q >> a.log_c) & (tmask >> 2);
      const unsigned g = q >> (a.log_c + a.log_t - 2);
      const unsigned t0 = ((u >> (log_h - 1)) << (log_h + 1)) | (u & (hh - 1));
      const unsigned i0 = (((g << a.log_t) | t0) << a.log_c) | c;
      const unsigned i1 = i0 + (hh << a.log_c), i2 = i0 + (h << a.log_c), i3 = i2 + (hh << a.log_c);
      const unsigned b = ((hi0 + g) << (s - a.s_lo)) | (t0 >> (a.s_hi - s));
      const fe *tw_s1 = a.tw + (2u << s) + 2 * b; // stage s+1 twiddles of blocks 2b, 2b+1
      fe x0 = tile[i0], x1 = tile[i1], x2 = tile[i2], x3 = tile[i3];
      if (!GS) {
        if (s != 0) {
          const fe wa = load_fe(a.tw + (1u << s) + b);
          x2 = mul_lazy<F>(x2, wa);
          x3 = mul_lazy<F>(x3, wa);
        }
        fe y0 = add2<F>(x0, x2), y2 = sub2<F>(x0, x2);
        fe y1 = add2<F>(x1, x3), y3 = sub2<F>(x1, x3);
        y1 = mul_lazy<F>(y1, load_fe(tw_s1));
        y3 = mul_lazy<F>(y3, load_fe(tw_s1 + 1));
        tile[i0] = add2<F>(y0, y1);
        tile[i1] = sub2<F>(y0, y1);
        tile[i2] = add2<F>(y2, y3);
        tile[i3] = sub2<F>(y2, y3);
      } else {
        const fe y0 = add2<F>(x0, x1), y1 = mul_lazy<F>(sub2<F>(x0, x1), load_fe(tw_s1));
        const fe y2 = add2<F>(x2, x3), y3 = mul_lazy<F>(sub2<F>(x2, x3), load_fe(tw_s1 + 1));
        fe z2 = sub2<F>(y0, y2), z3 = sub2<F>(y1, y3);
        if (s != 0) {
          const fe wa = load_fe(a.tw + (1u << s) + b);
          z2 = mul_lazy<F>(z2, wa);
          z3 = mul_lazy<F>(z3, wa);
        }
        tile[i0] = add2<F>(y0, y2);
        tile[i1] = add2<F>(y1, y3);
        tile[i2] = z2;
        tile[i3] = z3;
      }
    }
    __syncthreads();
  }

  for (unsigned e = threadIdx.x; e < elems; e += NTT_BLOCK) {
    const unsigned c = e & cmask, t = (e >> a.log_c) & tmask, g = e >> (a.log_c + a.log_t);
    const unsigned idx = ((hi0 + g) << (L - a.s_lo)) | (t << log_lo) | (lo0 + c);
    fe v = tile[e];
    switch (a.post) {
    case POST_REDUCE:
      v = reduce<F>(v);
      break;
    case POST_SCALE:
      v = mul<F>(v, a.scale);
      break;
    case POST_TABLE:
      v = mul<F>(v, twist_pow(a, GS ? idx : brev_bits(idx, L)));
      break;
    default:
      break;
    }
    store_fe(a.out + idx, v);
  }
}

// fine[j] = alpha^j (j < 2^12), coarse[j] = scale * alpha^(j << 12), so that
// twist(i) = fine[i mod 2^12] * coarse[i >> 12] = scale * alpha^i (fine[0] is skipped)
__global__ void twist_table_kernel(fe *fine, fe *coarse, fe alpha, fe alpha_4096, fe scale, unsigned coarse_count) {
  const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned nf = 1u << TWIST_FINE_BITS;
  if (j < nf)
    store_fe(fine + j, pow_u64<FrParams>(alpha, j));
  else if (j - nf < coarse_count)
    store_fe(coarse + (j - nf), mul<FrParams>(scale, pow_u64<FrParams>(alpha_4096, j - nf)));
}

struct pass_plan {
  unsigned s_lo, s_hi;
};

std::vector<pass_plan> plan_passes(unsigned L) {
  std::vector<pass_plan> p;
  const unsigned passes = (L + NTT_LOG_TILE - 1) / NTT_LOG_TILE;
  unsigned s = 0;
  for (unsigned i = 0; i < passes; i++) {
    const unsigned len = L / passes + (i < L % passes ? 1 : 0);
    p.push_back({s, s + len});
    s += len;
  }
  return p;
}

hipError_t set_kernel_attributes() {
  const int d = device_index();
  if (g_attr_set[d])
    return hipSuccess;
  const int bytes = (int)(sizeof(fe) << NTT_LOG_TILE);
  BC_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&ntt_pass_kernel<false>), hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  BC_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&ntt_pass_kernel<true>), hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
  g_attr_set[d] = true;
  return hipSuccess;
}

} // namespace

// Runs the whole transform on device buffers (in may equal out).
hipError_t ntt_device(const fe *in, fe *out, unsigned L, bool bitrev_in, bool inverse, unsigned log_ext, unsigned coset, hipMemPool_t pool,
                      hipStream_t stream) {
  using F = FrParams;
  if (L == 0 || L > NTT_MAX_LOG || L + log_ext > NTT_MAX_LOG)
    return hipErrorInvalidValue;
  if (log_ext && coset >= (1u << log_ext))
    return hipErrorInvalidValue;
  BC_TRY(set_kernel_attributes());
  const fe *tw_f = nullptr, *tw_i = nullptr;
  BC_TRY(ensure_twiddles(L, stream, &tw_f, &tw_i));

  // twist table: forward coset -> alpha = rho; inverse -> n^-1 * rho^-i (or just n^-1)
  fe n_inv = fe_one<F>();
  {
    fe two_inv = bc::inverse<F>(add<F>(fe_one<F>(), fe_one<F>()));
    for (unsigned i = 0; i < L; i++)
      n_inv = mul<F>(n_inv, two_inv);
  }
  fe *twist = nullptr;
  const unsigned coarse_count = L > TWIST_FINE_BITS ? 1u << (L - TWIST_FINE_BITS) : 1u;
  if (log_ext) {
    fe seven = fe_zero();
    seven.v[0] = F::generator;
    const fe g = to_mont<F>(seven);
    const fe w_big = pow_u64<F>(fe_const<F>(F::omega), 1ull << (F::omega_log_order - L - log_ext));
    fe rho = mul<F>(g, pow_u64<F>(w_big, coset));
    if (inverse)
      rho = bc::inverse<F>(rho);
    const fe rho_4096 = pow_u64<F>(rho, 1ull << TWIST_FINE_BITS);
    BC_TRY(pool_alloc(&twist, (1u << TWIST_FINE_BITS) + coarse_count, pool, stream));
    const unsigned threads = (1u << TWIST_FINE_BITS) + coarse_count;
    twist_table_kernel<<<grid_for(threads, 256), 256, 0, stream>>>(twist, twist + (1u << TWIST_FINE_BITS), rho, rho_4096, inverse ? n_inv : fe_one<F>(),
                                                                   coarse_count);
    BC_LAUNCH_CHECK();
  }

  const std::vector<pass_plan> plan = plan_passes(L);
  const unsigned np = (unsigned)plan.size();
  for (unsigned k = 0; k < np; k++) {
    const pass_plan pp = bitrev_in ? plan[np - 1 - k] : plan[k];
    pass_args a{};
    a.in = k == 0 ? in : out;
    a.out = out;
    a.tw = inverse ? tw_i : tw_f;
    a.log_n = L;
    a.s_lo = pp.s_lo;
    a.s_hi = pp.s_hi;
    a.log_t = pp.s_hi - pp.s_lo;
    const unsigned log_lo = L - pp.s_hi, log_hi = pp.s_lo;
    const unsigned room = NTT_LOG_TILE - a.log_t;
    a.log_c = log_lo < room ? log_lo : room;
    a.log_g = a.log_c == log_lo ? ((room - a.log_c) < log_hi ? room - a.log_c : log_hi) : 0;
    a.tiles = 1u << (L - a.log_t - a.log_c - a.log_g);
    a.xcd_remap = (a.tiles % 8 == 0) ? 1u : 0u;
    a.pre_table = k == 0 && log_ext && !inverse;
    a.twist_fine = twist;
    a.twist_coarse = twist ? twist + (1u << TWIST_FINE_BITS) : nullptr;
    if (k == np - 1) {
      if (!inverse)
        a.post = POST_REDUCE;
      else if (log_ext)
        a.post = POST_TABLE;
      else {
        a.post = POST_SCALE;
        a.scale = n_inv;
      }
    } else {
      a.post = POST_LAZY;
    }
    const size_t lds = sizeof(fe) << (a.log_g + a.log_t + a.log_c);
    if (bitrev_in)
      ntt_pass_kernel<true><<<a.tiles, NTT_BLOCK, lds, stream>>>(a);
    else
      ntt_pass_kernel<false><<<a.tiles, NTT_BLOCK, lds, stream>>>(a);
    BC_LAUNCH_CHECK();
  }
  if (twist)
    BC_TRY(pool_free(twist, stream));
  return hipSuccess;
}

hipError_t ntt_tear_down_impl() {
  std::lock_guard<std::mutex> lock(g_tw_mutex);
  twiddle_state &t = g_tw[device_index()];
  BC_TRY(hipDeviceSynchronize());
  for (fe *p : t.retired)
    BC_TRY(hipFree(p));
  if (t.fwd) {
    BC_TRY(hipFree(t.fwd));
    BC_TRY(hipFree(t.inv));
  }
  t = twiddle_state{};
  return hipSuccess;
}

hipError_t ntt_set_up_impl() {
  if (!ff_current())
    return hipErrorNotInitialized;
  BC_TRY(set_kernel_attributes());
  const fe *f, *i;
  BC_TRY(ensure_twiddles(20, nullptr, &f, &i));
  return hipDeviceSynchronize();
}

// Host-pointer handling of ntt_execute_async (src/ntt.cu:155-231): inputs and
// outputs may each be device, pinned or pageable memory.
hipError_t ntt_execute(const ntt_configuration &c) {
  hipStream_t s = to_stream(c.stream);
  hipMemPool_t pool = to_pool(c.mem_pool);
  const unsigned L = c.log_values_count;
  if (L == 0 || L > NTT_MAX_LOG)
    return hipErrorInvalidValue;
  const size_t bytes = sizeof(fe) << L;
  const bool in_host = classify_pointer(c.inputs) != mem_kind::device;
  const bool out_host = classify_pointer(c.outputs) != mem_kind::device;
  fe *d_in = static_cast<fe *>(c.inputs);
  fe *temp = nullptr;
  if (in_host) {
    BC_TRY(pool_alloc(&temp, (size_t)1 << L, pool, s));
    BC_TRY(hipMemcpyAsync(temp, c.inputs, bytes, hipMemcpyHostToDevice, s));
    if (c.h2d_copy_finished.handle)
      BC_TRY(hipEventRecord(to_event(c.h2d_copy_finished), s));
    if (c.h2d_copy_finished_callback)
      BC_TRY(hipLaunchHostFunc(s, c.h2d_copy_finished_callback, c.h2d_copy_finished_callback_data));
    d_in = temp;
  }
  fe *d_out = static_cast<fe *>(c.outputs);
  if (out_host) {
    if (in_host || c.can_overwrite_inputs)
      d_out = d_in;
    else {
      BC_TRY(pool_alloc(&temp, (size_t)1 << L, pool, s));
      d_out = temp;
    }
  }
  BC_TRY(ntt_device(d_in, d_out, L, c.bit_reversed_inputs, c.inverse, c.log_extension_degree, c.coset_index, pool, s));
  if (out_host) {
    BC_TRY(hipMemcpyAsync(c.outputs, d_out, bytes, hipMemcpyDeviceToHost, s));
    if (c.d2h_copy_finished.handle)
      BC_TRY(hipEventRecord(to_event(c.d2h_copy_finished), s));
    if (c.d2h_copy_finished_callback)
      BC_TRY(hipLaunchHostFunc(s, c.d2h_copy_finished_callback, c.d2h_copy_finished_callback_data));
  }
  if (temp)
    BC_TRY(pool_free(temp, s));
  return hipSuccess;
}

} // namespace bc

using namespace bc;

extern "C" {

bc_error ntt_set_up() { return as_bc(ntt_set_up_impl()); }

bc_error ntt_execute_async(ntt_configuration configuration) { return as_bc(ntt_execute(configuration)); }

bc_error ntt_execute_async_multigpu(const ntt_configuration *configurations, const int *dev_ids, unsigned log_n_devs) {
  if (log_n_devs != 0)
    return bc_error_invalid_value;
  int prev = 0;
  BC_TRY_BC(hipGetDevice(&prev));
  BC_TRY_BC(hipSetDevice(dev_ids[0]));
  const hipError_t e = ntt_execute(configurations[0]);
  (void)hipSetDevice(prev);
  return as_bc(e);
}

bc_error ntt_tear_down() { return as_bc(ntt_tear_down_impl()); }

} // extern "C"
