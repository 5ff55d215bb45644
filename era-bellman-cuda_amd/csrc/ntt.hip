// Fr NTT for gfx950, batched: a batch of transforms of any sizes and
// directions runs as one launch per pass index (per butterfly order), with
// the per-transform descriptors passed by value in the kernel arguments.  The
// plain C-ABI call is a batch of one.
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
// run of at most NTT_LOG_TILE consecutive stages [s_lo, s_hi) on LDS tiles of
// 2^NTT_LOG_TILE = 1024 elements (32 KiB, five blocks per CU): the stages'
// butterflies stay inside groups of 2^(s_hi-s_lo) elements, strided by
// 2^(L-s_hi), so a transform of 2^L takes ceil(L / 10) HBM passes (three at
// 2^24).  Tiles are mapped XCD-major so that neighbouring strided columns run
// on the same XCD and share its L2 lines.
#include <algorithm>
#include <mutex>
#include <vector>

#include "batch.h"
#include "common.h"
#include "ntt.h"
#include "tables.h"
#include "tuning.h"

namespace bc {

namespace {

constexpr unsigned NTT_BLOCK = (1u << NTT_LOG_TILE) / 4; // one radix-4 group per thread and step
constexpr unsigned TWIST_FINE_BITS = 12;

struct twiddle_state {
  fe *fwd = nullptr;       // TW[2^s + b] = omega_{2^(s+1)}^brev_s(b), index 0 unused
  fe *inv = nullptr;       // inverse powers, same layout
  fe *inv_sizes = nullptr; // 2^-i, i <= NTT_MAX_LOG
  unsigned log_size = 0;
  bool attributes = false;
  std::vector<fe *> retired;
};

twiddle_state g_tw[MAX_DEVICES];
std::mutex g_tw_mutex;

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

enum : uint8_t { POST_LAZY = 0, POST_REDUCE = 1, POST_SCALE = 2, POST_TABLE = 3 };

// One transform's share of a pass launch.
struct pass_job {
  const fe *in;
  fe *out;
  const fe *aux; // POST_SCALE: n^-1; twists: fine table, coarse table 2^12 entries later
  uint32_t tile_begin; // first block of the job (a multiple of 8)
  uint8_t log_n, s_lo, s_hi;
  uint8_t log_c, log_g; // tile = 2^log_g groups x 2^(s_hi-s_lo) rows x 2^log_c columns
  uint8_t inverse, pre_table, post;
};

struct pass_table {
  uint32_t count;
  const fe *tw_fwd;
  const fe *tw_inv;
  pass_job job[NTT_MAX_JOBS];
};
static_assert(sizeof(pass_table) <= 4096, "kernel argument budget");

__device__ __forceinline__ fe twist_pow(const fe *fine, unsigned i) {
  const fe c = load_fe(fine + (1u << TWIST_FINE_BITS) + (i >> TWIST_FINE_BITS));
  const unsigned f = i & ((1u << TWIST_FINE_BITS) - 1);
  return f ? mul<FrParams>(load_fe(fine + f), c) : c;
}

template <bool GS> __global__ __launch_bounds__(NTT_BLOCK) void ntt_pass_kernel(const pass_table pt) {
  using F = FrParams;
  extern __shared__ fe tile[];
  unsigned jl = 0, jh = pt.count; // job of this block
  while (jh - jl > 1) {
    const unsigned mid = (jl + jh) >> 1;
    if (pt.job[mid].tile_begin <= blockIdx.x)
      jl = mid;
    else
      jh = mid;
  }
  const pass_job &a = pt.job[jl];
  const unsigned L = a.log_n, s_lo = a.s_lo, s_hi = a.s_hi, log_c = a.log_c, log_g = a.log_g;
  const unsigned log_t = s_hi - s_lo;
  const unsigned tiles = 1u << (L - log_t - log_c - log_g);
  unsigned tid_tile = blockIdx.x - a.tile_begin;
  if (tid_tile >= tiles)
    return; // alignment padding
  if ((tiles & 7) == 0)
    tid_tile = (tid_tile & 7) * (tiles >> 3) + (tid_tile >> 3);
  const fe *__restrict__ tw = a.inverse ? pt.tw_inv : pt.tw_fwd;
  const unsigned log_lo = L - s_hi;             // bits below the group
  const unsigned lo_tiles_log = log_lo - log_c; // tiles along the low bits
  const unsigned lo0 = (tid_tile & ((1u << lo_tiles_log) - 1)) << log_c;
  const unsigned hi0 = (tid_tile >> lo_tiles_log) << log_g;
  const unsigned elems = 1u << (log_g + log_t + log_c);
  const unsigned cmask = (1u << log_c) - 1, tmask = (1u << log_t) - 1;

  // element e of the tile <-> index of the transform
  auto index_of = [&](unsigned e) {
    const unsigned c = e & cmask, t = (e >> log_c) & tmask, g = e >> (log_c + log_t);
    return ((hi0 + g) << (L - s_lo)) | (t << log_lo) | (lo0 + c);
  };
  constexpr unsigned PER_THREAD = (1u << NTT_LOG_TILE) / NTT_BLOCK;
  if (a.pre_table && elems == (1u << NTT_LOG_TILE)) {
    // coset twist: the elements and their twists are loaded first, then the
    // products are formed
    fe v[PER_THREAD], tw[PER_THREAD];
#pragma unroll
    for (unsigned r = 0; r < PER_THREAD; r++) {
      const unsigned idx = index_of(threadIdx.x + r * NTT_BLOCK);
      v[r] = load_fe(a.in + idx);
      tw[r] = twist_pow(a.aux, GS ? brev_bits(idx, L) : idx);
    }
#pragma unroll
    for (unsigned r = 0; r < PER_THREAD; r++)
      tile[threadIdx.x + r * NTT_BLOCK] = mul_lazy<F>(v[r], tw[r]);
  } else if (a.pre_table) {
    for (unsigned e = threadIdx.x; e < elems; e += NTT_BLOCK) {
      const unsigned idx = index_of(e);
      tile[e] = mul_lazy<F>(load_fe(a.in + idx), twist_pow(a.aux, GS ? brev_bits(idx, L) : idx));
    }
  } else if (elems == (1u << NTT_LOG_TILE)) {
    // a thread's loads are all issued before the first is used
    fe v[PER_THREAD];
#pragma unroll
    for (unsigned r = 0; r < PER_THREAD; r++)
      v[r] = load_fe(a.in + index_of(threadIdx.x + r * NTT_BLOCK));
#pragma unroll
    for (unsigned r = 0; r < PER_THREAD; r++)
      tile[threadIdx.x + r * NTT_BLOCK] = v[r];
  } else {
    for (unsigned e = threadIdx.x; e < elems; e += NTT_BLOCK)
      tile[e] = load_fe(a.in + index_of(e));
  }
  __syncthreads();

  // Stages in network order (CT: s_lo..s_hi-1, GS: s_hi-1..s_lo).  An odd
  // count starts with one radix-2 step; the rest go two at a time: a thread
  // holds the 4 elements t0, t0+h/2, t0+h, t0+3h/2 (h = row distance of the
  // lower stage s) and applies stages s and s+1 in registers.
  const unsigned nst = log_t;
  unsigned done = 0;
  if (nst & 1) {
    const unsigned s = GS ? s_hi - 1 : s_lo;
    const unsigned log_h = s_hi - s - 1;
    const unsigned h = 1u << log_h;
    for (unsigned q = threadIdx.x; q < (elems >> 1); q += NTT_BLOCK) {
      const unsigned c = q & cmask;
      const unsigned r = (q >> log_c) & (tmask >> 1);
      const unsigned g = q >> (log_c + log_t - 1);
      const unsigned t0 = ((r >> log_h) << (log_h + 1)) | (r & (h - 1));
      const unsigned i0 = (((g << log_t) | t0) << log_c) | c;
      const unsigned i1 = i0 + (h << log_c);
      const unsigned b = ((hi0 + g) << (s - s_lo)) | (t0 >> (s_hi - s));
      const fe E = tile[i0], O = tile[i1];
      if (!GS) {
        // CT tile values are < 4m; each output is < 4m again
        const fe wo = s != 0 ? mul_lazy<F>(O, load_fe(tw + (1u << s) + b)) : reduce2<F>(O);
        const fe e = reduce2<F>(E);
        tile[i0] = add_raw(e, wo);
        tile[i1] = sub_raw2<F>(e, wo);
      } else {
        // GS tile values stay < 2m
        tile[i0] = add2<F>(E, O);
        tile[i1] = s != 0 ? mul_lazy<F>(sub_raw2<F>(E, O), load_fe(tw + (1u << s) + b)) : sub2<F>(E, O);
      }
    }
    __syncthreads();
    done = 1;
  }
  for (; done < nst; done += 2) {
    // the lower-numbered stage of the pair
    const unsigned s = GS ? s_hi - 2 - done : s_lo + done;
    const unsigned log_h = s_hi - s - 1; // >= 1
    const unsigned h = 1u << log_h, hh = h >> 1;
    for (unsigned q = threadIdx.x; q < (elems >> 2); q += NTT_BLOCK) {
      const unsigned c = q & cmask;
      const unsigned u = (q >> log_c) & (tmask >> 2);
      const unsigned g = q >> (log_c + log_t - 2);
      const unsigned t0 = ((u >> (log_h - 1)) << (log_h + 1)) | (u & (hh - 1));
      const unsigned i0 = (((g << log_t) | t0) << log_c) | c;
      const unsigned i1 = i0 + (hh << log_c), i2 = i0 + (h << log_c), i3 = i2 + (hh << log_c);
      const unsigned b = ((hi0 + g) << (s - s_lo)) | (t0 >> (s_hi - s));
      const fe *tw_s1 = tw + (2u << s) + 2 * b; // stage s+1 twiddles of blocks 2b, 2b+1
      // the step's twiddles are loaded ahead of the tile reads (for s = 0,
      // TW[1] = 1 is loaded and not used)
      const fe wa = load_fe(tw + (1u << s) + b), wb0 = load_fe(tw_s1), wb1 = load_fe(tw_s1 + 1);
      fe x0 = tile[i0], x1 = tile[i1], x2 = tile[i2], x3 = tile[i3];
      if (!GS) {
        // values < 4m in and out: the odd inputs of a butterfly go through a
        // twiddle product (< 2m), the even ones through one reduction (< 2m)
        if (s != 0) {
          x2 = mul_lazy<F>(x2, wa);
          x3 = mul_lazy<F>(x3, wa);
        } else {
          x2 = reduce2<F>(x2);
          x3 = reduce2<F>(x3);
        }
        x0 = reduce2<F>(x0);
        x1 = reduce2<F>(x1);
        const fe y0 = reduce2<F>(add_raw(x0, x2)), y2 = reduce2<F>(sub_raw2<F>(x0, x2));
        const fe y1 = mul_lazy<F>(add_raw(x1, x3), wb0);
        const fe y3 = mul_lazy<F>(sub_raw2<F>(x1, x3), wb1);
        tile[i0] = add_raw(y0, y1);
        tile[i1] = sub_raw2<F>(y0, y1);
        tile[i2] = add_raw(y2, y3);
        tile[i3] = sub_raw2<F>(y2, y3);
      } else {
        // values < 2m in and out; differences that feed a product skip the reduction
        const fe y0 = add2<F>(x0, x1), y1 = mul_lazy<F>(sub_raw2<F>(x0, x1), wb0);
        const fe y2 = add2<F>(x2, x3), y3 = mul_lazy<F>(sub_raw2<F>(x2, x3), wb1);
        fe z2, z3;
        if (s != 0) {
          z2 = mul_lazy<F>(sub_raw2<F>(y0, y2), wa);
          z3 = mul_lazy<F>(sub_raw2<F>(y1, y3), wa);
        } else {
          z2 = sub2<F>(y0, y2);
          z3 = sub2<F>(y1, y3);
        }
        tile[i0] = add2<F>(y0, y2);
        tile[i1] = add2<F>(y1, y3);
        tile[i2] = z2;
        tile[i3] = z3;
      }
    }
    __syncthreads();
  }

  const unsigned post = a.post;
  const fe scale = post == POST_SCALE ? load_fe(a.aux) : fe_zero();
  for (unsigned e = threadIdx.x; e < elems; e += NTT_BLOCK) {
    const unsigned c = e & cmask, t = (e >> log_c) & tmask, g = e >> (log_c + log_t);
    const unsigned idx = ((hi0 + g) << (L - s_lo)) | (t << log_lo) | (lo0 + c);
    fe v = tile[e];
    switch (post) {
    case POST_REDUCE:
      v = reduce4<F>(v); // CT tiles hold values < 4m
      break;
    case POST_SCALE:
      v = mul<F>(v, scale);
      break;
    case POST_TABLE:
      v = mul<F>(v, twist_pow(a.aux, GS ? idx : brev_bits(idx, L)));
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

// Stage runs of the HBM passes: ceil(L / NTT_LOG_TILE) passes, balanced;
// the tuning table may name the first pass's stage count, and the rest are
// then balanced over the remaining passes.
std::vector<pass_plan> plan_passes(unsigned L) {
  std::vector<pass_plan> p;
  unsigned passes = ntt_passes(L), s = 0;
  const unsigned first = passes >= 2 ? ntt_first_pass(L) : 0;
  if (first && ntt_split_valid(L, first)) {
    p.push_back({0, first});
    s = first;
    passes--;
  }
  const unsigned rest = L - s;
  for (unsigned i = 0; i < passes; i++) {
    const unsigned len = rest / passes + (i < rest % passes ? 1 : 0);
    p.push_back({s, s + len});
    s += len;
  }
  return p;
}

fe inverse_power_of_two(unsigned L) {
  using F = FrParams;
  const fe two_inv = bc::inverse<F>(add<F>(fe_one<F>(), fe_one<F>()));
  fe r = fe_one<F>();
  for (unsigned i = 0; i < L; i++)
    r = mul<F>(r, two_inv);
  return r;
}

} // namespace

// Twiddles for 2^max_log_n and kernel attributes of the current device.  Growing
// the tables synchronises `stream`; older tables stay alive until ntt_tear_down
// because in-flight work may still read them.
hipError_t ntt_prepare(unsigned max_log_n, hipStream_t stream, ntt_tables *out) {
  std::lock_guard<std::mutex> lock(g_tw_mutex);
  const int d = device_slot();
  if (d < 0)
    return hipErrorInvalidDevice;
  twiddle_state &t = g_tw[d];
  if (!t.attributes) {
    const int bytes = (int)(sizeof(fe) << NTT_LOG_TILE);
    BC_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&ntt_pass_kernel<false>), hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    BC_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&ntt_pass_kernel<true>), hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    t.attributes = true;
  }
  if (!t.inv_sizes) {
    std::vector<fe> sizes(NTT_MAX_LOG + 1);
    for (unsigned i = 0; i <= NTT_MAX_LOG; i++)
      sizes[i] = inverse_power_of_two(i);
    fe *d = nullptr;
    BC_TRY(hipMalloc(&d, sizeof(fe) * sizes.size()));
    BC_TRY(hipMemcpy(d, sizes.data(), sizeof(fe) * sizes.size(), hipMemcpyHostToDevice));
    t.inv_sizes = d;
  }
  if (t.log_size < max_log_n) {
    const ff_state *ff = ff_current();
    if (!ff)
      return hipErrorNotInitialized;
    const unsigned log_size = std::max(max_log_n, 12u);
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
  if (out)
    *out = ntt_tables{t.fwd, t.inv, t.inv_sizes, t.log_size};
  return hipSuccess;
}

hipError_t ntt_batch_device(const ntt_job *jobs, unsigned count, arena &ar) {
  using F = FrParams;
  hipStream_t stream = ar.stream();
  if (count == 0)
    return hipSuccess;
  if (count > NTT_MAX_JOBS)
    return hipErrorInvalidValue;
  unsigned max_log = 0;
  for (unsigned j = 0; j < count; j++) {
    const ntt_job &q = jobs[j];
    if (!ntt_job_valid(q))
      return hipErrorInvalidValue;
    max_log = std::max(max_log, q.log_n);
  }
  ntt_tables tt{};
  BC_TRY(ntt_prepare(max_log, stream, &tt));

  // per-job twist tables: forward coset -> rho^i; inverse -> n^-1 * rho^-i
  std::vector<fe *> twist(count, nullptr);
  for (unsigned j = 0; j < count; j++) {
    const ntt_job &q = jobs[j];
    if (!q.log_ext)
      continue;
    const unsigned L = q.log_n;
    const unsigned coarse_count = L > TWIST_FINE_BITS ? 1u << (L - TWIST_FINE_BITS) : 1u;
    fe seven = fe_zero();
    seven.v[0] = F::generator;
    const fe g = to_mont<F>(seven);
    const fe w_big = pow_u64<F>(fe_const<F>(F::omega), 1ull << (F::omega_log_order - L - q.log_ext));
    fe rho = mul<F>(g, pow_u64<F>(w_big, q.coset));
    if (q.inverse)
      rho = bc::inverse<F>(rho);
    const fe rho_4096 = pow_u64<F>(rho, 1ull << TWIST_FINE_BITS);
    const unsigned entries = (1u << TWIST_FINE_BITS) + coarse_count;
    BC_TRY(ar.get(&twist[j], entries));
    twist_table_kernel<<<grid_for(entries, 256), 256, 0, stream>>>(twist[j], twist[j] + (1u << TWIST_FINE_BITS), rho, rho_4096,
                                                                  q.inverse ? inverse_power_of_two(L) : fe_one<F>(), coarse_count);
    BC_LAUNCH_CHECK();
  }

  std::vector<std::vector<pass_plan>> plans(count);
  unsigned max_passes = 0;
  for (unsigned j = 0; j < count; j++) {
    plans[j] = plan_passes(jobs[j].log_n);
    max_passes = std::max(max_passes, (unsigned)plans[j].size());
  }
  for (unsigned k = 0; k < max_passes; k++) {
    for (int gs = 0; gs < 2; gs++) {
      pass_table pt{};
      pt.tw_fwd = tt.fwd;
      pt.tw_inv = tt.inv;
      uint32_t blocks = 0;
      size_t lds = 0;
      for (unsigned j = 0; j < count; j++) {
        const ntt_job &q = jobs[j];
        const unsigned np = (unsigned)plans[j].size();
        if ((int)q.bitrev_in != gs || k >= np)
          continue;
        const pass_plan pp = q.bitrev_in ? plans[j][np - 1 - k] : plans[j][k];
        const unsigned L = q.log_n;
        const unsigned log_t = pp.s_hi - pp.s_lo;
        const unsigned log_lo = L - pp.s_hi, log_hi = pp.s_lo;
        const unsigned room = NTT_LOG_TILE - log_t;
        const unsigned log_c = std::min(log_lo, room);
        const unsigned log_g = log_c == log_lo ? std::min(room - log_c, log_hi) : 0;
        pass_job &a = pt.job[pt.count++];
        a.in = k == 0 ? q.in : q.out;
        a.out = q.out;
        a.log_n = (uint8_t)L;
        a.s_lo = (uint8_t)pp.s_lo;
        a.s_hi = (uint8_t)pp.s_hi;
        a.log_c = (uint8_t)log_c;
        a.log_g = (uint8_t)log_g;
        a.inverse = q.inverse ? 1 : 0;
        a.pre_table = (k == 0 && q.log_ext && !q.inverse) ? 1 : 0;
        a.aux = twist[j];
        if (k == np - 1) {
          if (!q.inverse)
            a.post = POST_REDUCE;
          else if (q.log_ext)
            a.post = POST_TABLE;
          else {
            a.post = POST_SCALE;
            a.aux = tt.inv_sizes + L;
          }
        } else {
          a.post = POST_LAZY;
        }
        a.tile_begin = blocks;
        const uint32_t tiles = 1u << (L - log_t - log_c - log_g);
        blocks += (tiles + 7) & ~7u;
        lds = std::max(lds, sizeof(fe) << (log_g + log_t + log_c));
      }
      if (pt.count == 0)
        continue;
      if (gs)
        ntt_pass_kernel<true><<<blocks, NTT_BLOCK, lds, stream>>>(pt);
      else
        ntt_pass_kernel<false><<<blocks, NTT_BLOCK, lds, stream>>>(pt);
      BC_LAUNCH_CHECK();
    }
  }
  for (fe *p : twist)
    if (p)
      BC_TRY(ar.put(p));
  return hipSuccess;
}

hipError_t ntt_tear_down_impl() {
  std::lock_guard<std::mutex> lock(g_tw_mutex);
  const int d = device_slot();
  if (d < 0)
    return hipErrorInvalidDevice;
  twiddle_state &t = g_tw[d];
  BC_TRY(hipDeviceSynchronize());
  for (fe *p : t.retired)
    BC_TRY(hipFree(p));
  if (t.fwd) {
    BC_TRY(hipFree(t.fwd));
    BC_TRY(hipFree(t.inv));
  }
  if (t.inv_sizes)
    BC_TRY(hipFree(t.inv_sizes));
  t = twiddle_state{};
  return hipSuccess;
}

hipError_t ntt_set_up_impl() {
  if (!ff_current())
    return hipErrorNotInitialized;
  BC_TRY(ntt_prepare(20, nullptr, nullptr));
  return hipDeviceSynchronize();
}

// Host-pointer handling of ntt_execute_async (src/ntt.cu:155-231): inputs and
// outputs may each be device, pinned or pageable memory.  Temporaries are
// released on every return path.
hipError_t ntt_execute(const ntt_configuration &c) {
  hipStream_t s = to_stream(c.stream);
  const unsigned L = c.log_values_count;
  // checked before anything is allocated or copied
  if (!ntt_shape_valid(L, c.log_extension_degree, c.coset_index) || !c.inputs || !c.outputs)
    return hipErrorInvalidValue;
  call_resources res(to_pool(c.mem_pool), s);
  const size_t bytes = sizeof(fe) << L;
  const bool in_host = classify_pointer(c.inputs) != mem_kind::device;
  const bool out_host = classify_pointer(c.outputs) != mem_kind::device;
  fe *d_in = static_cast<fe *>(c.inputs);
  fe *d_out = static_cast<fe *>(c.outputs);
  fe *staged_in = nullptr;
  if (in_host)
    BC_TRY(res.get(&staged_in, (size_t)1 << L));
  if (out_host && !in_host && !c.can_overwrite_inputs)
    BC_TRY(res.get(&d_out, (size_t)1 << L));
  const ntt_job probe{d_in, d_out, L, c.bit_reversed_inputs, c.inverse, c.log_extension_degree, c.coset_index};
  const size_t block_bytes = ntt_footprint(probe);
  char *block = nullptr;
  BC_TRY(res.get(&block, block_bytes));
  if (in_host) {
    BC_TRY(hipMemcpyAsync(staged_in, c.inputs, bytes, hipMemcpyHostToDevice, s));
    if (c.h2d_copy_finished.handle)
      BC_TRY(hipEventRecord(to_event(c.h2d_copy_finished), s));
    if (c.h2d_copy_finished_callback)
      BC_TRY(hipLaunchHostFunc(s, c.h2d_copy_finished_callback, c.h2d_copy_finished_callback_data));
    d_in = staged_in;
  }
  if (out_host && (in_host || c.can_overwrite_inputs))
    d_out = d_in; // transform in place in the staging copy (or the caller's device inputs)
  const ntt_job job{d_in, d_out, L, c.bit_reversed_inputs, c.inverse, c.log_extension_degree, c.coset_index};
  arena ar = arena::bump_mode(block, block_bytes, s);
  BC_TRY(ntt_batch_device(&job, 1, ar));
  if (out_host) {
    BC_TRY(hipMemcpyAsync(c.outputs, d_out, bytes, hipMemcpyDeviceToHost, s));
    if (c.d2h_copy_finished.handle)
      BC_TRY(hipEventRecord(to_event(c.d2h_copy_finished), s));
    if (c.d2h_copy_finished_callback)
      BC_TRY(hipLaunchHostFunc(s, c.d2h_copy_finished_callback, c.d2h_copy_finished_callback_data));
  }
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
