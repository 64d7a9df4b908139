// Wave-resident pass kernel for gfx950 (CDNA4, wave64).
//
// Each wave owns a "wave tile" of 1024 amplitudes of the group-local state.
// For the gates the tile is held in the wave layout: tile bits 0..5 are the
// lane (one complex128 per lane, a 1 KiB coalesced row) and tile bits 6..9 are
// four register slots (16 complex128 per lane in VGPRs) placed on the pass's
// register qubits E[0..3].  Gates on lane bits exchange partners across lanes,
// gates on register slots are in-lane, and runs of diagonal gates become
// per-lane and per-slot phase factors.
//
// The codec runs in the chunk layout: one LDS transpose gives lane L the tile
// positions 16L..16L+15, i.e. one half of codec chunk L>>1 (32 consecutive
// amplitudes, 64 doubles).  Exponent and width reductions are then in-register
// plus one partner-lane DPP step, each lane's codes form one contiguous bit
// range of its chunk's payload, and the tile's payload is staged in LDS and
// written with one allocation atomic and a coalesced copy.  There is no
// work-group barrier: waves progress independently.  Results are bit-identical
// to the LDS-tile kernel (same chunk format and quantiser).
#include <utility>

#include "qsim_device.h"

namespace qsim {

constexpr int kRegBits = 4;
constexpr int kV = 1 << kRegBits;  // complex values per lane

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
  const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, l);
  const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 shfl_xor2(double2 v, int m) {
  return make_double2(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64));
}
// value of lane ^ (1 << B) without the LDS crossbar: DPP within a row of 16
// lanes (B <= 3), the row-swap instructions for B = 4 and 5
template <int B>
__device__ __forceinline__ uint32_t lane_xor_u32(uint32_t v, uint32_t lane) {
  static_assert(B >= 0 && B < 6, "lane bit");
  if constexpr (B == 0) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xb1, 0xf, 0xf, false);  // quad_perm [1,0,3,2]
  } else if constexpr (B == 1) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4e, 0xf, 0xf, false);  // quad_perm [2,3,0,1]
  } else if constexpr (B == 2) {
    // row_half_mirror (i ^ 7), then quad_perm [3,2,1,0] (i ^ 3)
    const int h = __builtin_amdgcn_mov_dpp((int)v, 0x141, 0xf, 0xf, false);
    return (uint32_t)__builtin_amdgcn_mov_dpp(h, 0x1b, 0xf, 0xf, false);
  } else if constexpr (B == 3) {
    // row_mirror (i ^ 15), then row_half_mirror (i ^ 7)
    const int h = __builtin_amdgcn_mov_dpp((int)v, 0x140, 0xf, 0xf, false);
    return (uint32_t)__builtin_amdgcn_mov_dpp(h, 0x141, 0xf, 0xf, false);
  } else if constexpr (B == 4) {
    // with both operands v: the first result holds rows (0, 0, 2, 2), the
    // second rows (1, 1, 3, 3)
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (lane & 16u) ? r[0] : r[1];
  } else {
    // first result: halves (lo, lo); second: (hi, hi)
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (lane & 32u) ? r[0] : r[1];
  }
}
template <int B>
__device__ __forceinline__ double lane_xor_f64(double v, uint32_t lane) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint64_t r = ((uint64_t)lane_xor_u32<B>((uint32_t)(b >> 32), lane) << 32) | lane_xor_u32<B>((uint32_t)b, lane);
  return __longlong_as_double((long long)r);
}
template <int B>
__device__ __forceinline__ double2 lane_xor2(double2 v, uint32_t lane) {
  return make_double2(lane_xor_f64<B>(v.x, lane), lane_xor_f64<B>(v.y, lane));
}
// For B = 4 or 5: exchanges the lanes of x whose bit B is set with the lanes
// of y whose bit B is clear.  Afterwards every lane holds in x the amplitude
// with bit B clear and in y its partner with bit B set (of row x in the lanes
// with bit B clear, of row y in the others); a second call restores the rows.
template <int B>
__device__ __forceinline__ void swap_half_rows(double2& x, double2& y) {
  static_assert(B == 4 || B == 5, "row swaps exist for lane bits 4 and 5");
  uint64_t xb[2] = {(uint64_t)__double_as_longlong(x.x), (uint64_t)__double_as_longlong(x.y)};
  uint64_t yb[2] = {(uint64_t)__double_as_longlong(y.x), (uint64_t)__double_as_longlong(y.y)};
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    uint32_t xl = (uint32_t)xb[c], xh = (uint32_t)(xb[c] >> 32), yl = (uint32_t)yb[c], yh = (uint32_t)(yb[c] >> 32);
    if constexpr (B == 4) {
      const auto rl = __builtin_amdgcn_permlane16_swap(xl, yl, false, false);
      const auto rh = __builtin_amdgcn_permlane16_swap(xh, yh, false, false);
      xl = rl[0]; yl = rl[1]; xh = rh[0]; yh = rh[1];
    } else {
      const auto rl = __builtin_amdgcn_permlane32_swap(xl, yl, false, false);
      const auto rh = __builtin_amdgcn_permlane32_swap(xh, yh, false, false);
      xl = rl[0]; yl = rl[1]; xh = rh[0]; yh = rh[1];
    }
    xb[c] = ((uint64_t)xh << 32) | xl;
    yb[c] = ((uint64_t)yh << 32) | yl;
  }
  x = make_double2(__longlong_as_double((long long)xb[0]), __longlong_as_double((long long)xb[1]));
  y = make_double2(__longlong_as_double((long long)yb[0]), __longlong_as_double((long long)yb[1]));
}

// keeps the compiler from hoisting coefficient loads of later register
// iterations (bounds VGPR live ranges)
#define QSIM_ITER_FENCE() asm volatile("" ::: "memory")

// one PhaseTerm as eight dwords (a single scalar load)
typedef uint32_t TermBits __attribute__((ext_vector_type(8)));
static_assert(sizeof(PhaseTerm) == sizeof(TermBits), "PhaseTerm layout");

__device__ __forceinline__ double2 ldc(const QC double* __restrict__ c, uint64_t idx) {
  return make_double2(c[2 * idx], c[2 * idx + 1]);
}
// out = m0*a + m1*b
__device__ __forceinline__ double2 mac2(double2 m0, double2 a, double2 m1, double2 b) {
  double re = 0.0, im = 0.0;
  cmad(re, im, m0.x, m0.y, a.x, a.y);
  cmad(re, im, m1.x, m1.y, b.x, b.y);
  return make_double2(re, im);
}
// out = m0*a + m1*b for real m0, m1.  The products and sums are those of
// mac2 without its imaginary-part terms; with a zero imaginary part those
// terms add a signed zero to a sum that is never -0 (an fma onto +0 or onto a
// sum of that kind), so both give the same bits.
__device__ __forceinline__ double2 mac2r(double m0, double2 a, double m1, double2 b) {
  double re = fma(m0, a.x, 0.0), im = fma(m0, a.y, 0.0);
  re = fma(m1, b.x, re);
  im = fma(m1, b.y, im);
  return make_double2(re, im);
}
// out = d*a + (i o)*b and out = (i o)*a + d*b for real d, o, in mac2's order
// of products (the same bits, by the argument above)
__device__ __forceinline__ double2 mac2_ri(double d, double2 a, double o, double2 b) {
  double re = fma(d, a.x, 0.0), im = fma(d, a.y, 0.0);
  re = fma(-o, b.y, re);
  im = fma(o, b.x, im);
  return make_double2(re, im);
}
__device__ __forceinline__ double2 mac2_ir(double o, double2 a, double d, double2 b) {
  double re = fma(-o, a.y, 0.0), im = fma(o, a.x, 0.0);
  re = fma(d, b.x, re);
  im = fma(d, b.y, im);
  return make_double2(re, im);
}

// ------------------------------------------------------------------ codec
// value of the partner lane (lane ^ 1)
__device__ __forceinline__ uint32_t swap1_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xb1, 0xf, 0xf, false);  // quad_perm [1,0,3,2]
}
__device__ __forceinline__ double swap1_f64(double v) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint64_t r = ((uint64_t)swap1_u32((uint32_t)(b >> 32)) << 32) | swap1_u32((uint32_t)b);
  return __longlong_as_double((long long)r);
}
__device__ __forceinline__ int swap1_i32(int v) { return (int)swap1_u32((uint32_t)v); }

// min(a, |b|) with the IEEE NaN rule (a NaN operand yields the other one)
__device__ __forceinline__ double vmin_abs(double a, double b) {
  double r;
  asm("v_min_f64 %0, %1, |%2|" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// wave-wide inclusive prefix sum (gfx9 DPP row shifts and row broadcasts)
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

// LDS slot of tile position p for the layout transposes: slot p + (p >> PS),
// PS = log2(KV).  Wave-layout accesses (64v + l) become slot RS v + l + (l >>
// PS) with RS = 64 + 64/KV, chunk-layout accesses (KV L + q, q < KV) slot
// (KV + 1) L + q: a per-lane base plus an immediate offset, and free of bank
// conflicts in both directions (the eight lanes of each 128-bit access group
// touch eight distinct 16-byte bank groups).
template <int KV>
constexpr uint32_t x_pad_shift() { return KV == 16 ? 4u : 3u; }
template <int KV>
constexpr uint32_t x_row_slots() { return 64u + (64u >> x_pad_shift<KV>()); }
template <int KV>
constexpr uint32_t x_slots() { return x_row_slots<KV>() * KV; }

// wave layout (lane l, row v: tile position 64v + l) -> chunk layout
// (lane L, row q: tile position KV L + q), through the wave's LDS buffer
template <int KV>
__device__ __forceinline__ void to_chunk_layout(double2 (&amp)[KV], double2* xb, uint32_t lane) {
  constexpr uint32_t PS = x_pad_shift<KV>(), RS = x_row_slots<KV>();
  double2* wa = xb + lane + (lane >> PS);
#pragma unroll
  for (int v = 0; v < KV; ++v) wa[RS * v] = amp[v];
  __builtin_amdgcn_wave_barrier();
  const double2* ca = xb + (KV + 1) * lane;
#pragma unroll
  for (int q = 0; q < KV; ++q) amp[q] = ca[q];
  __builtin_amdgcn_wave_barrier();
}
// the same, with the tile permuted on the way: wave-layout value (l, v) goes
// to slot sl + srow[v] (the slot of its permuted position), and the chunk
// layout is read as above
template <int KV>
__device__ __forceinline__ void to_chunk_layout_perm(double2 (&amp)[KV], double2* xb, uint32_t lane, uint32_t sl,
                                                     const QC uint16_t* __restrict__ srow) {
  double2* wa = xb + sl;
#pragma unroll
  for (int v = 0; v < KV; ++v) wa[srow[v]] = amp[v];
  __builtin_amdgcn_wave_barrier();
  const double2* ca = xb + (KV + 1) * lane;
#pragma unroll
  for (int q = 0; q < KV; ++q) amp[q] = ca[q];
  __builtin_amdgcn_wave_barrier();
}
template <int KV>
__device__ __forceinline__ void to_wave_layout(double2 (&amp)[KV], double2* xb, uint32_t lane) {
  constexpr uint32_t PS = x_pad_shift<KV>(), RS = x_row_slots<KV>();
  double2* ca = xb + (KV + 1) * lane;
#pragma unroll
  for (int q = 0; q < KV; ++q) ca[q] = amp[q];
  __builtin_amdgcn_wave_barrier();
  const double2* wa = xb + lane + (lane >> PS);
#pragma unroll
  for (int v = 0; v < KV; ++v) amp[v] = wa[RS * v];
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ double unzig_ldexp(uint64_t u, int e) {
  const int64_t k = (int64_t)(u >> 1) ^ -(int64_t)(u & 1ull);
  return ldexp((double)k, e);
}
__device__ __forceinline__ uint64_t extract64(const QG uint64_t* __restrict__ P, uint32_t bit, uint32_t w) {
  const uint32_t q = bit >> 6, s = bit & 63u;
  uint64_t v = P[q] >> s;
  if (s + w > 64u) v |= P[q + 1] << (64u - s);
  return v & ((1ull << w) - 1ull);
}

// Double di of a chunk, any width 0..64 (the general path, branch-free):
// the code is read from the two payload words that can hold it (the arena
// keeps 64 bytes of padding past its last payload, so both are mapped).
__device__ __forceinline__ double decode_at(uint64_t d, const QG uint64_t* __restrict__ arena, uint32_t di) {
  const uint32_t w = desc_w(d);
  const QG uint64_t* P = arena + desc_off(d);
  const uint32_t bit = di * w, qw = bit >> 6, sh = bit & 63u;
  const uint64_t v = (P[qw] >> sh) | ((P[qw + 1] << 1) << (63u - sh));
  const uint64_t u = w >= 64u ? v : v & ((1ull << w) - 1ull);
  const double x = w >= 64u ? __longlong_as_double((long long)u) : unzig_ldexp(u, desc_e(d));
  return w == 0u ? 0.0 : x;
}

// OR a code of w <= 64 bits at bit offset `bit` into an LDS staging row (the
// second OR is zero when the field does not cross a word boundary)
__device__ __forceinline__ void stage_or(uint64_t* stg, uint32_t bit, uint64_t c) {
  const uint32_t qw = bit >> 6, sh = bit & 63u;
  atomicOr(reinterpret_cast<unsigned long long*>(&stg[qw]), (unsigned long long)(c << sh));
  atomicOr(reinterpret_cast<unsigned long long*>(&stg[qw + 1]), (unsigned long long)((c >> 1) >> (63u - sh)));
}

// Pack this lane's 2 KV codes of w <= 32 bits (chunk layout: doubles d0 ..
// d0 + 2 KV - 1 of its chunk, code of double d at bit d w) into the staging
// row, M consecutive codes per OR (M w <= 64).  An OR that would add nothing
// (an all-zero piece, or a piece that ends inside its first word) goes to
// the lane's own spare word `dummy` instead: lanes of one chunk then never
// OR into one word when their codes are zero (w = 0 chunks, whose lanes all
// start at the same bit), and narrow codes take 2 KV / M ORs instead of KV.
template <int KV, int M>
__device__ __forceinline__ void pack_pieces(uint64_t* sc, uint64_t* dummy, const double2 (&amp)[KV], uint32_t d0,
                                            uint32_t w) {
  static_assert((2 * KV) % M == 0, "pieces tile the lane's codes");
#pragma unroll
  for (int p = 0; p < 2 * KV / M; ++p) {
    uint64_t F = 0;
#pragma unroll
    for (int i = 0; i < M; ++i) {
      const int d = p * M + i;
      const uint32_t c = (uint32_t)__double_as_longlong((d & 1) ? amp[d >> 1].y : amp[d >> 1].x);
      F |= i == 0 ? (uint64_t)c : (uint64_t)c << ((uint32_t)i * w);
    }
    const uint32_t bit = (d0 + (uint32_t)(p * M)) * w, qw = bit >> 6, sh = bit & 63u;
    uint64_t* a0 = F != 0ull ? sc + qw : dummy;
    uint64_t* a1 = F != 0ull && sh + (uint32_t)M * w > 64u ? sc + qw + 1 : dummy;
    atomicOr(reinterpret_cast<unsigned long long*>(a0), (unsigned long long)(F << sh));
    atomicOr(reinterpret_cast<unsigned long long*>(a1), (unsigned long long)(sh ? F >> (64u - sh) : 0ull));
  }
}

// Quantise the part of a chunk held by this lane (chunk layout: a chunk
// spans 32/KV lanes, two for KV = 16 and four for KV = 8) and replace amp by
// the codes (zig-zag integers of rint(x * 2^-e), or the raw bit patterns when
// the chunk is stored verbatim).  Returns the chunk's w | (e & 0xffff) << 8,
// the same in every lane of the chunk.
//
// e is the minimum over the chunk's nonzero doubles of
// ilogb(2*fma(eps_rel,|x|,eps_abs)) minus the guard bits.  That function is
// non-decreasing in |x|, so it is evaluated once, at the smallest nonzero |x|
// (found as min over |bits(x) - 1| read as doubles: zeros become NaN patterns,
// which v_min_f64 skips).  The largest high word bounds |x| exactly against
// powers of two, which decides the 2^51 verbatim bound and the 32-bit code path.
// A high word at or above ovf_hi (non-finite values, or values whose error
// budget overflows) takes the exact per-double exponent path.
// rint is the 1.5*2^52 shift (round-to-nearest-even for |s| < 2^51).
// Every lane of the wave must call it.
// value of lane ^ 2 (quad_perm [2,3,0,1])
__device__ __forceinline__ uint32_t swap2_u32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4e, 0xf, 0xf, false);
}
__device__ __forceinline__ double swap2_f64(double v) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint64_t r = ((uint64_t)swap2_u32((uint32_t)(b >> 32)) << 32) | swap2_u32((uint32_t)b);
  return __longlong_as_double((long long)r);
}
// reductions over the 32/KV lanes that share a chunk (lane ^ 1, and lane ^ 2
// when a chunk spans four lanes)
template <int KV>
__device__ __forceinline__ double chunk_fmin(double x) {
  x = fmin(x, swap1_f64(x));
  if constexpr (KV == 8) x = fmin(x, swap2_f64(x));
  return x;
}
template <int KV>
__device__ __forceinline__ uint32_t chunk_umax(uint32_t x) {
  x = max(x, swap1_u32(x));
  if constexpr (KV == 8) x = max(x, swap2_u32(x));
  return x;
}
template <int KV>
__device__ __forceinline__ int chunk_imin(int x) {
  x = min(x, swap1_i32(x));
  if constexpr (KV == 8) x = min(x, (int)swap2_u32((uint32_t)x));
  return x;
}
template <int KV>
__device__ __forceinline__ uint32_t chunk_uor(uint32_t x) {
  x |= swap1_u32(x);
  if constexpr (KV == 8) x |= swap2_u32(x);
  return x;
}

template <int KV>
__device__ __forceinline__ uint32_t lane_quantise(double2 (&amp)[KV], const CodecParams& cp, uint32_t ovf_hi) {
  static_assert(KV == 16 || KV == 8, "a chunk spans two or four lanes");
  double mn = __longlong_as_double(0x7ff0000000000000ll);  // +inf
  uint32_t hmax = 0;
#pragma unroll
  for (int q = 0; q < KV; ++q) {
    const uint64_t bx = (uint64_t)__double_as_longlong(amp[q].x), by = (uint64_t)__double_as_longlong(amp[q].y);
    mn = vmin_abs(mn, __longlong_as_double((long long)(bx - 1ull)));
    mn = vmin_abs(mn, __longlong_as_double((long long)(by - 1ull)));
    hmax = max(hmax, max((uint32_t)(bx >> 32) & 0x7fffffffu, (uint32_t)(by >> 32) & 0x7fffffffu));
  }
  mn = chunk_fmin<KV>(mn);
  hmax = chunk_umax<KV>(hmax);
  int le;
  if (hmax >= ovf_hi) {
    int l2 = 0x7fffffff;
#pragma unroll
    for (int q = 0; q < KV; ++q) l2 = min(l2, min(lane_exponent(amp[q].x, cp), lane_exponent(amp[q].y, cp)));
    le = chunk_imin<KV>(l2);
  } else {
    const bool nzc = (uint64_t)__double_as_longlong(mn) != 0x7ff0000000000000ull;
    le = nzc ? lane_exponent(__longlong_as_double(__double_as_longlong(mn) + 1ll), cp) : 0x7fffffff;
  }
  if (le == 0x7fffffff) {
    // all-zero chunk: no payload; its codes must be 0 (a -0.0 bit pattern
    // would otherwise be OR-ed into the next chunk's first word)
#pragma unroll
    for (int q = 0; q < KV; ++q) amp[q] = make_double2(0.0, 0.0);
    return 0u;
  }
  bool raw = le == (int)0x80000000 || le < -1000;
  const int e = raw ? 0 : le - cp.guard;
  raw = raw || hmax >= ((uint32_t)(51 + 1023 + e) << 20);  // some |x * 2^-e| >= 2^51 (or non-finite)
  if (raw) return 64u;
  const double M = 6755399441055744.0;  // 1.5 * 2^52
  // x * 2^-e is exact here (|x * 2^-e| < 2^51 and 2^-e >= 2^-1074), so one
  // fma rounds x * 2^-e + M exactly as the scaled sum would
  const double sc = ldexp(1.0, -e);
  uint32_t w;
  if (hmax < ((uint32_t)(30 + 1023 + e) << 20)) {
    // every |x * 2^-e| < 2^30: 32-bit codes; the low dword of the shifted
    // double is the integer in two's complement
    uint32_t orr = 0;
#pragma unroll
    for (int q = 0; q < KV; ++q) {
      const int k0 = (int)(uint32_t)__double_as_longlong(fma(amp[q].x, sc, M));
      const int k1 = (int)(uint32_t)__double_as_longlong(fma(amp[q].y, sc, M));
      const uint32_t u0 = ((uint32_t)k0 << 1) ^ (uint32_t)(k0 >> 31), u1 = ((uint32_t)k1 << 1) ^ (uint32_t)(k1 >> 31);
      orr |= u0 | u1;
      amp[q] = make_double2(__longlong_as_double((long long)u0), __longlong_as_double((long long)u1));
    }
    orr = chunk_uor<KV>(orr);
    w = orr ? 32u - (uint32_t)__clz((int)orr) : 0u;
  } else {
    uint64_t orr = 0;
#pragma unroll
    for (int q = 0; q < KV; ++q) {
      const int64_t k0 = __double_as_longlong(fma(amp[q].x, sc, M)) - 0x4338000000000000ll;
      const int64_t k1 = __double_as_longlong(fma(amp[q].y, sc, M)) - 0x4338000000000000ll;
      const uint64_t u0 = ((uint64_t)k0 << 1) ^ (uint64_t)(k0 >> 63), u1 = ((uint64_t)k1 << 1) ^ (uint64_t)(k1 >> 63);
      orr |= u0 | u1;
      amp[q] = make_double2(__longlong_as_double((long long)u0), __longlong_as_double((long long)u1));
    }
    orr = ((uint64_t)chunk_uor<KV>((uint32_t)(orr >> 32)) << 32) | chunk_uor<KV>((uint32_t)orr);
    w = orr ? 64u - (uint32_t)__clzll(orr) : 0u;
  }
  return w | ((uint32_t)(e & 0xffff) << 8);
}

// ------------------------------------------------------------ dense ops
struct WaveCtx {
  int lane;
  uint64_t base;
  uint64_t tau;          // the tile's index in its group (phase tables)
  const double* ptab;    // the pass's phase tables
};

// selector of the op for register index v: bits from lane, slots or base
__device__ __forceinline__ uint32_t wave_sel(const QC Op& op, const WaveCtx& cx, int v) {
  uint32_t sel = 0;
  for (uint32_t s = 0; s < op.nsel; ++s) {
    const uint32_t pos = op.sel_pos[s];
    uint32_t bit;
    if (!op.sel_tile[s]) bit = (uint32_t)((cx.base >> pos) & 1ull);
    else if (pos < 6) bit = ((uint32_t)cx.lane >> pos) & 1u;
    else bit = ((uint32_t)v >> (pos - 6)) & 1u;
    sel |= bit << s;
  }
  return sel;
}

// true when no selector bit comes from a lane: the selector is the same on
// every lane of a row
__device__ __forceinline__ bool op_sel_uniform(const QC Op& op) {
  for (uint32_t s = 0; s < op.nsel; ++s)
    if (op.sel_tile[s] && op.sel_pos[s] < 6) return false;
  return true;
}

// one target on lane bit B, one matrix (at complex offset mo) for every row
template <int B, int KV>
__device__ __forceinline__ void k1_lane_uniform(const QC Op& op, double2 (&amp)[KV], const WaveCtx& cx,
                                                const QC double* __restrict__ C, uint64_t mo) {
  const uint32_t lane = (uint32_t)cx.lane;
  if constexpr (B >= 4) {
    // rows in pairs: after the half-row swap each lane holds both amplitudes
    // of a pair, so the matrix entries are wave-uniform
    const double2 m0 = ldc(C, mo), m1 = ldc(C, mo + 1), m2 = ldc(C, mo + 2), m3 = ldc(C, mo + 3);
    if (op.real == kDenseReal) {
#pragma unroll
      for (int v = 0; v < KV; v += 2) {
        double2 x = amp[v], y = amp[v + 1];
        swap_half_rows<B>(x, y);
        double2 x2 = mac2r(m0.x, x, m1.x, y), y2 = mac2r(m2.x, x, m3.x, y);
        swap_half_rows<B>(x2, y2);
        amp[v] = x2;
        amp[v + 1] = y2;
      }
      return;
    }
    if (op.real == kDenseRealDiagImagOff) {
#pragma unroll
      for (int v = 0; v < KV; v += 2) {
        double2 x = amp[v], y = amp[v + 1];
        swap_half_rows<B>(x, y);
        double2 x2 = mac2_ri(m0.x, x, m1.y, y), y2 = mac2_ir(m2.y, x, m3.x, y);
        swap_half_rows<B>(x2, y2);
        amp[v] = x2;
        amp[v + 1] = y2;
      }
      return;
    }
#pragma unroll
    for (int v = 0; v < KV; v += 2) {
      double2 x = amp[v], y = amp[v + 1];
      swap_half_rows<B>(x, y);
      double2 x2 = mac2(m0, x, m1, y), y2 = mac2(m2, x, m3, y);
      swap_half_rows<B>(x2, y2);
      amp[v] = x2;
      amp[v + 1] = y2;
    }
  } else {
    const uint32_t r = (lane >> B) & 1u;
    // the four entries are wave-uniform (scalar loads); each lane takes its
    // row's diagonal and off-diagonal entry
    const double2 m0 = ldc(C, mo), m1 = ldc(C, mo + 1), m2 = ldc(C, mo + 2), m3 = ldc(C, mo + 3);
    const double2 md = r ? m3 : m0, mf = r ? m2 : m1;
    if (op.real == kDenseReal) {
#pragma unroll
      for (int v = 0; v < KV; ++v) amp[v] = mac2r(md.x, amp[v], mf.x, lane_xor2<B>(amp[v], lane));
      return;
    }
    if (op.real == kDenseRealDiagImagOff) {
      // md is the diagonal entry of this lane's row, mf the off-diagonal one
#pragma unroll
      for (int v = 0; v < KV; ++v) amp[v] = mac2_ri(md.x, amp[v], mf.y, lane_xor2<B>(amp[v], lane));
      return;
    }
#pragma unroll
    for (int v = 0; v < KV; ++v) amp[v] = mac2(md, amp[v], mf, lane_xor2<B>(amp[v], lane));
  }
}

// one target on lane bit b, matrix chosen per element by the selector bits
template <int KV>
__device__ __forceinline__ void k1_lane_sel(const QC Op& op, double2 (&amp)[KV], const WaveCtx& cx,
                                            const QC double* __restrict__ C) {
  const uint32_t b = op.tpos[0];
  const uint32_t r = ((uint32_t)cx.lane >> b) & 1u;
  const int m = 1 << b;
#pragma unroll
  for (int v = 0; v < KV; ++v) {
    QSIM_ITER_FENCE();
    const double2 p = shfl_xor2(amp[v], m);
    const uint32_t sel = wave_sel(op, cx, v);
    if ((op.ident_mask >> sel) & 1ull) continue;
    const uint64_t mo0 = op.mat_off + 4ull * sel;
    amp[v] = mac2(ldc(C, mo0 + 3u * r), amp[v], ldc(C, mo0 + 1u + r), p);
  }
}

// one target on register slot J
template <int J, int KV>
__device__ __forceinline__ void k1_reg(const QC Op& op, double2 (&amp)[KV], const WaveCtx& cx, const QC double* __restrict__ C) {
  constexpr int MJ = 1 << J;
  if constexpr (MJ >= KV) return;  // no such register qubit in this tile shape
  if (op.nsel == 0) {
    // one matrix for every row: its entries are loaded once (wave-uniform)
    if (op.ident_mask & 1ull) return;
    const uint64_t mo = op.mat_off;
    const double2 m0 = ldc(C, mo), m1 = ldc(C, mo + 1), m2 = ldc(C, mo + 2), m3 = ldc(C, mo + 3);
    if (op.real == kDenseReal) {
#pragma unroll
      for (int v = 0; v < KV; ++v) {
        if (v & MJ) continue;
        const double2 a = amp[v], b = amp[v | MJ];
        amp[v] = mac2r(m0.x, a, m1.x, b);
        amp[v | MJ] = mac2r(m2.x, a, m3.x, b);
      }
      return;
    }
    if (op.real == kDenseRealDiagImagOff) {
#pragma unroll
      for (int v = 0; v < KV; ++v) {
        if (v & MJ) continue;
        const double2 a = amp[v], b = amp[v | MJ];
        amp[v] = mac2_ri(m0.x, a, m1.y, b);
        amp[v | MJ] = mac2_ir(m2.y, a, m3.x, b);
      }
      return;
    }
#pragma unroll
    for (int v = 0; v < KV; ++v) {
      if (v & MJ) continue;
      const double2 a = amp[v], b = amp[v | MJ];
      amp[v] = mac2(m0, a, m1, b);
      amp[v | MJ] = mac2(m2, a, m3, b);
    }
    return;
  }
  if (op_sel_uniform(op)) {
    // the selector is the same on every lane: the matrix comes through
    // scalar loads
#pragma unroll
    for (int v = 0; v < KV; ++v) {
      if (v & MJ) continue;
      QSIM_ITER_FENCE();
      const uint32_t sel = (uint32_t)__builtin_amdgcn_readfirstlane((int)wave_sel(op, cx, v));
      if ((op.ident_mask >> sel) & 1ull) continue;
      const uint64_t mo = op.mat_off + 4ull * sel;
      const double2 a = amp[v], b = amp[v | MJ];
      amp[v] = mac2(ldc(C, mo), a, ldc(C, mo + 1), b);
      amp[v | MJ] = mac2(ldc(C, mo + 2), a, ldc(C, mo + 3), b);
    }
    return;
  }
#pragma unroll
  for (int v = 0; v < KV; ++v) {
    if (v & MJ) continue;
    QSIM_ITER_FENCE();
    const uint32_t sel = wave_sel(op, cx, v);
    if ((op.ident_mask >> sel) & 1ull) continue;
    const uint64_t mo = op.mat_off + 4ull * sel;
    const double2 a = amp[v], b = amp[v | MJ];
    amp[v] = mac2(ldc(C, mo), a, ldc(C, mo + 1), b);
    amp[v | MJ] = mac2(ldc(C, mo + 2), a, ldc(C, mo + 3), b);
  }
}

// the 4x4 product on rows v, v | M0, v | M1, v | M0 | M1 (matrix at mo)
template <int M0, int M1, int KV>
__device__ __forceinline__ void k2_regreg_rows(double2 (&amp)[KV], int v, const QC double* __restrict__ C, uint64_t mo) {
  {
    const double2 x[4] = {amp[v], amp[v | M0], amp[v | M1], amp[v | M0 | M1]};
    double2 y[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double re = 0.0, im = 0.0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double2 m = ldc(C, mo + 4 * r + c);
        cmad(re, im, m.x, m.y, x[c].x, x[c].y);
      }
      y[r] = make_double2(re, im);
    }
    amp[v] = y[0];
    amp[v | M0] = y[1];
    amp[v | M1] = y[2];
    amp[v | M0 | M1] = y[3];
  }
}

// two targets on register slots J0 (matrix bit 0) and J1 (matrix bit 1)
template <int J0, int J1, int KV>
__device__ __forceinline__ void k2_regreg(const QC Op& op, double2 (&amp)[KV], const WaveCtx& cx, const QC double* __restrict__ C) {
  constexpr int M0 = 1 << J0, M1 = 1 << J1;
  if constexpr (M0 >= KV || M1 >= KV) return;  // no such register qubit in this tile shape
  // a wave-uniform selector keeps the matrix in scalar loads (the uniform
  // form is a separate loop, so that the compiler sees uniform addresses)
  if (op.nsel == 0 || op_sel_uniform(op)) {
#pragma unroll
    for (int v = 0; v < KV; ++v) {
      if (v & (M0 | M1)) continue;
      QSIM_ITER_FENCE();
      const uint32_t sel = (uint32_t)__builtin_amdgcn_readfirstlane((int)wave_sel(op, cx, v));
      if ((op.ident_mask >> sel) & 1ull) continue;
      k2_regreg_rows<M0, M1, KV>(amp, v, C, op.mat_off + 16ull * sel);
    }
    return;
  }
#pragma unroll
  for (int v = 0; v < KV; ++v) {
    if (v & (M0 | M1)) continue;
    QSIM_ITER_FENCE();
    const uint32_t sel = wave_sel(op, cx, v);
    if ((op.ident_mask >> sel) & 1ull) continue;
    k2_regreg_rows<M0, M1, KV>(amp, v, C, op.mat_off + 16ull * sel);
  }
}
// two targets: lane bit b (matrix bit il) and register slot J (matrix bit 1-il)
template <int J, int KV>
__device__ __forceinline__ void k2_lanereg(const QC Op& op, double2 (&amp)[KV], const WaveCtx& cx, const QC double* __restrict__ C,
                                           uint32_t b, uint32_t il) {
  constexpr int MJ = 1 << J;
  if constexpr (MJ >= KV) return;  // no such register qubit in this tile shape
  const uint32_t ir = 1u - il;
  const uint32_t rl = ((uint32_t)cx.lane >> b) & 1u;
  const int m = 1 << b;
#pragma unroll
  for (int v = 0; v < KV; ++v) {
    if (v & MJ) continue;
    QSIM_ITER_FENCE();
    const double2 a0 = amp[v], a1 = amp[v | MJ];
    const double2 p0 = shfl_xor2(a0, m), p1 = shfl_xor2(a1, m);
    const uint32_t sel = op.nsel ? wave_sel(op, cx, v) : 0u;
    if ((op.ident_mask >> sel) & 1ull) continue;
    const uint64_t mo = op.mat_off + 16ull * sel;
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const uint32_t row = (rl << il) | ((uint32_t)g << ir);
      double re = 0.0, im = 0.0;
#pragma unroll
      for (int l = 0; l < 2; ++l)
#pragma unroll
        for (int gc = 0; gc < 2; ++gc) {
          // column (lane bit l, register bit gc): own lane holds lane bit rl
          const double2 val = ((uint32_t)l == rl) ? (gc ? a1 : a0) : (gc ? p1 : p0);
          const double2 mm = ldc(C, mo + 4ull * row + (((uint32_t)l << il) | ((uint32_t)gc << ir)));
          cmad(re, im, mm.x, mm.y, val.x, val.y);
        }
      if (g == 0) amp[v] = make_double2(re, im);
      else amp[v | MJ] = make_double2(re, im);
    }
  }
}

// two targets on lane bits b0 (matrix bit 0) and b1 (matrix bit 1)
template <int KV>
__device__ __forceinline__ void k2_lanelane(const QC Op& op, double2 (&amp)[KV], const WaveCtx& cx, const QC double* __restrict__ C) {
  const uint32_t b0 = op.tpos[0], b1 = op.tpos[1];
  const uint32_t r = (((uint32_t)cx.lane >> b0) & 1u) | ((((uint32_t)cx.lane >> b1) & 1u) << 1);
#pragma unroll
  for (int v = 0; v < KV; ++v) {
    QSIM_ITER_FENCE();
    const double2 x1 = shfl_xor2(amp[v], 1 << b0);
    const double2 x2 = shfl_xor2(amp[v], 1 << b1);
    const double2 x3 = shfl_xor2(amp[v], (1 << b0) | (1 << b1));
    const uint32_t sel = op.nsel ? wave_sel(op, cx, v) : 0u;
    if ((op.ident_mask >> sel) & 1ull) continue;
    const uint64_t mo = op.mat_off + 16ull * sel + 4ull * r;
    double re = 0.0, im = 0.0;
    const double2 own = amp[v];
    const double2 vals[4] = {own, x1, x2, x3};  // indexed by column ^ r
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      const double2 mm = ldc(C, mo + (r ^ (uint32_t)f));
      cmad(re, im, mm.x, mm.y, vals[f].x, vals[f].y);
    }
    amp[v] = make_double2(re, im);
  }
}

template <int KV>
__device__ __forceinline__ void wave_dense(const QC Op& op, double2 (&amp)[KV], const WaveCtx& cx, const QC double* __restrict__ C) {
  if (op.k == 1) {
    const uint32_t t = op.tpos[0];
    if (t < 6) {
      uint64_t mo = op.mat_off;
      if (op.nsel != 0) {
        bool tile_sel = false;
        for (uint32_t i = 0; i < op.nsel; ++i) tile_sel = tile_sel || op.sel_tile[i];
        if (tile_sel) {
          // a selector on a lane or a register slot: the matrix varies per element
          k1_lane_sel(op, amp, cx, C);
          return;
        }
        // selectors on base bits only: one matrix for the whole tile
        const uint32_t sel = (uint32_t)__builtin_amdgcn_readfirstlane((int)wave_sel(op, cx, 0));
        if ((op.ident_mask >> sel) & 1ull) return;
        mo += 4ull * sel;
      } else if (op.ident_mask & 1ull) {
        return;
      }
      switch (t) {
        case 0: k1_lane_uniform<0>(op, amp, cx, C, mo); break;
        case 1: k1_lane_uniform<1>(op, amp, cx, C, mo); break;
        case 2: k1_lane_uniform<2>(op, amp, cx, C, mo); break;
        case 3: k1_lane_uniform<3>(op, amp, cx, C, mo); break;
        case 4: k1_lane_uniform<4>(op, amp, cx, C, mo); break;
        default: k1_lane_uniform<5>(op, amp, cx, C, mo); break;
      }
    } else switch (t - 6) {
        case 0: k1_reg<0>(op, amp, cx, C); break;
        case 1: k1_reg<1>(op, amp, cx, C); break;
        case 2: k1_reg<2>(op, amp, cx, C); break;
        default: k1_reg<3>(op, amp, cx, C); break;
      }
    return;
  }
  const uint32_t t0 = op.tpos[0], t1 = op.tpos[1];
  if (t0 < 6 && t1 < 6) { k2_lanelane(op, amp, cx, C); return; }
  if (t0 >= 6 && t1 >= 6) {
    const uint32_t code = (t0 - 6) * 4 + (t1 - 6);
    switch (code) {
      case 1: k2_regreg<0, 1>(op, amp, cx, C); break;
      case 2: k2_regreg<0, 2>(op, amp, cx, C); break;
      case 3: k2_regreg<0, 3>(op, amp, cx, C); break;
      case 4: k2_regreg<1, 0>(op, amp, cx, C); break;
      case 6: k2_regreg<1, 2>(op, amp, cx, C); break;
      case 7: k2_regreg<1, 3>(op, amp, cx, C); break;
      case 8: k2_regreg<2, 0>(op, amp, cx, C); break;
      case 9: k2_regreg<2, 1>(op, amp, cx, C); break;
      case 11: k2_regreg<2, 3>(op, amp, cx, C); break;
      case 12: k2_regreg<3, 0>(op, amp, cx, C); break;
      case 13: k2_regreg<3, 1>(op, amp, cx, C); break;
      case 14: k2_regreg<3, 2>(op, amp, cx, C); break;
      default: break;
    }
    return;
  }
  const uint32_t b = t0 < 6 ? t0 : t1, j = (t0 < 6 ? t1 : t0) - 6, il = t0 < 6 ? 0u : 1u;
  switch (j) {
    case 0: k2_lanereg<0>(op, amp, cx, C, b, il); break;
    case 1: k2_lanereg<1>(op, amp, cx, C, b, il); break;
    case 2: k2_lanereg<2>(op, amp, cx, C, b, il); break;
    default: k2_lanereg<3>(op, amp, cx, C, b, il); break;
  }
}

// Run of diagonal gates.  Terms conditioned on lane bits (and base bits) fold
// into one per-lane factor; terms that also require one register slot bit fold
// into a per-lane factor of that slot; the element's phase is the product of
// the lane factor and the factors of its set slot bits.  Terms on two or more
// slots are summed per element.
template <int KV>
__device__ __forceinline__ void wave_elem(const QC Op& op, double2 (&amp)[KV], const WaveCtx& cx,
                                          const QC PhaseTerm* __restrict__ terms, const QC DiagTable* __restrict__ diags,
                                          const QC double* __restrict__ C) {
  const uint32_t lane = (uint32_t)cx.lane;
  if (op.ptab_off != kNoPhaseTable) {
    // phase tables: a per-lane factor of the tile, per slot a per-lane
    // factor times a per-tile one, and per row a uniform factor
    const QG double* TG = (const QG double*)(cx.ptab) + 2 * op.ptab_off;
    const QC double* TU = (const QC double*)(cx.ptab) + 2 * op.ptab_off;
    auto ldg = [&](uint64_t i) { return make_double2(TG[2 * i], TG[2 * i + 1]); };
    const uint32_t sl = op.ptab_slots, fl = op.ptab_flags, gb = (sl >> 8) & 63u;
    // the slot-free factor: Q[lane] M[row][lane] U[tau] prod_k G[tau][k]^(lane bit k),
    // M's row the tile (kPtabM) or the value of the key bits of the base (kPtabK)
    if ((fl & (kPtabQ | kPtabU | kPtabM | kPtabK)) || gb) {
      double2 t0 = make_double2(1.0, 0.0);
      if (fl & kPtabQ) t0 = ldg(op.ptab_q + lane);
      if (fl & (kPtabM | kPtabK)) {
        uint64_t row = cx.tau;
        if (fl & kPtabK) {
          row = 0;
          for (uint32_t i = 0; i < op.ptab_nk; ++i) row |= ((cx.base >> op.ptab_kpos[i]) & 1ull) << i;
        }
        const double2 m = ldg(op.ptab_m + (row << 6) + lane);
        t0 = (fl & kPtabQ) ? cmul(t0, m) : m;
      }
      if (fl & kPtabU) t0 = cmul(t0, ldc(TU, op.ptab_q + 64 + cx.tau));
      if (gb) {
        // G holds angles: the lane's factor is cis of the sum over its set bits
        const QC double* G = TU + 2 * (op.ptab_g + (cx.tau << 3));
        double ga = 0.0;
#pragma unroll
        for (uint32_t k = 0; k < 6; ++k)
          if (gb & (1u << k)) ga += ((lane >> k) & 1u) ? G[2 * k] : 0.0;
        t0 = cmul(t0, cis_reduced(ga));
      }
#pragma unroll
      for (int v = 0; v < KV; ++v) amp[v] = cmul(amp[v], t0);
    }
    // slot factors, one at a time (each table read right before its use
    // keeps the registers of one factor live)
    const uint64_t F = kPtabF + (cx.tau << 2);
    if (sl & 1u) {
      __builtin_amdgcn_sched_barrier(0);
      const double2 f = (fl & kPtabFt) ? cmul(ldg(kPtabC + lane), ldc(TU, F + 0)) : ldg(kPtabC + lane);
#pragma unroll
      for (int v = 1; v < KV; v += 2) amp[v] = cmul(amp[v], f);
    }
    if (sl & 2u) {
      __builtin_amdgcn_sched_barrier(0);
      const double2 f = (fl & kPtabFt) ? cmul(ldg(kPtabC + 64 + lane), ldc(TU, F + 1)) : ldg(kPtabC + 64 + lane);
#pragma unroll
      for (int v = 0; v < KV; ++v)
        if (v & 2) amp[v] = cmul(amp[v], f);
    }
    if (sl & 4u) {
      __builtin_amdgcn_sched_barrier(0);
      const double2 f = (fl & kPtabFt) ? cmul(ldg(kPtabC + 128 + lane), ldc(TU, F + 2)) : ldg(kPtabC + 128 + lane);
#pragma unroll
      for (int v = 0; v < KV; ++v)
        if (v & 4) amp[v] = cmul(amp[v], f);
    }
    if (KV > 8 && (sl & 8u)) {
      __builtin_amdgcn_sched_barrier(0);
      const double2 f = (fl & kPtabFt) ? cmul(ldg(kPtabC + 192 + lane), ldc(TU, F + 3)) : ldg(kPtabC + 192 + lane);
#pragma unroll
      for (int v = 8; v < KV; ++v) amp[v] = cmul(amp[v], f);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (fl & kPtabRows) {
#pragma unroll
      for (int v = 0; v < KV; ++v) amp[v] = cmul(amp[v], ldc(TU, kPtabR + v));
    }
  } else if (op.nterms) {
    double aL = 0.0, aF0 = 0.0, aF1 = 0.0, aF2 = 0.0, aF3 = 0.0;
    uint32_t used = 0;
    bool generic = false;
    // kTermGroup terms per scalar load round trip (each op's list is padded
    // to whole groups and the array ends with one spare group), the next
    // group requested before the current one is applied
    const QC TermBits* tp = (const QC TermBits*)(terms + op.term_off);
    TermBits tg[kTermGroup];
#pragma unroll
    for (uint32_t j = 0; j < kTermGroup; ++j) tg[j] = tp[j];
    for (uint32_t i0 = 0; i0 < op.nterms; i0 += kTermGroup) {
      TermBits tn[kTermGroup];
#pragma unroll
      for (uint32_t j = 0; j < kTermGroup; ++j) tn[j] = tp[i0 + kTermGroup + j];
#pragma unroll
      for (uint32_t j = 0; j < kTermGroup; ++j) {
      const uint32_t tmask = tg[j][0], tval = tg[j][1];
      const uint64_t bmask = (uint64_t)tg[j][2] | ((uint64_t)tg[j][3] << 32);
      const uint64_t bval = (uint64_t)tg[j][4] | ((uint64_t)tg[j][5] << 32);
      if ((cx.base & bmask) != bval) continue;
      const double angle = __longlong_as_double((long long)((uint64_t)tg[j][6] | ((uint64_t)tg[j][7] << 32)));
      const uint32_t lm = tmask & 63u, lv = tval & 63u;
      const uint32_t rm = (tmask >> 6) & 15u, rv = (tval >> 6) & 15u;
      const double a = ((lane & lm) == lv) ? angle : 0.0;
      if (rm == 0) { aL += a; used |= 16u; continue; }
      if (rm & (rm - 1u)) { generic = true; continue; }
      const uint32_t js = __ffs(rm) - 1;
      double f = a;
      if (!rv) { aL += a; f = -a; used |= 16u; }
      aF0 += js == 0 ? f : 0.0;
      aF1 += js == 1 ? f : 0.0;
      aF2 += js == 2 ? f : 0.0;
      aF3 += js == 3 ? f : 0.0;
      used |= 1u << js;
      }
#pragma unroll
      for (uint32_t j = 0; j < kTermGroup; ++j) tg[j] = tn[j];
    }
    // each factor is formed right before it is applied; the scheduling
    // barriers keep the compiler from evaluating all five at once
    if (used & 16u) {
      const double2 F = cis_reduced(aL);
#pragma unroll
      for (int v = 0; v < KV; ++v) amp[v] = cmul(amp[v], F);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (used & 1u) {
      const double2 F = cis_reduced(aF0);
#pragma unroll
      for (int v = 1; v < KV; v += 2) amp[v] = cmul(amp[v], F);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (used & 2u) {
      const double2 F = cis_reduced(aF1);
#pragma unroll
      for (int v = 0; v < KV; ++v)
        if (v & 2) amp[v] = cmul(amp[v], F);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (used & 4u) {
      const double2 F = cis_reduced(aF2);
#pragma unroll
      for (int v = 0; v < KV; ++v)
        if (v & 4) amp[v] = cmul(amp[v], F);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (used & 8u) {
      const double2 F = cis_reduced(aF3);
#pragma unroll
      for (int v = 8; v < KV; ++v) amp[v] = cmul(amp[v], F);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (generic) {
      // terms on two or more slots: one uniform phase per term, applied to
      // the register elements whose slot bits match and the lanes whose lane
      // bits match
      for (uint32_t i = 0; i < op.nterms; ++i) {
        const QC PhaseTerm& T = terms[op.term_off + i];
        const uint32_t rm = (T.tmask >> 6) & 15u, rv = (T.tval >> 6) & 15u;
        if (!(rm & (rm - 1u))) continue;
        if ((cx.base & T.bmask) != T.bval) continue;
        const uint32_t lm = T.tmask & 63u, lv = T.tval & 63u;
        const double2 f = ((lane & lm) == lv) ? cis_reduced(T.angle) : make_double2(1.0, 0.0);
#pragma unroll
        for (int v = 0; v < KV; ++v)
          if (((uint32_t)v & rm) == rv) {
            // the row test is wave-uniform: the fence keeps it a scalar
            // branch, so only the matching rows (a quarter or less) pay the
            // multiply instead of every row paying it and a select
            QSIM_ITER_FENCE();
            amp[v] = cmul(amp[v], f);
          }
      }
    }
  }
  for (uint32_t d = 0; d < op.ndiag; ++d) {
    const QC DiagTable& Dt = diags[op.diag_off + d];
    // index bits from base and lane are fixed for the wave; slot bits vary by row
    uint32_t idx0 = 0, slot_pos = 0, slot_bit[kRegBits] = {0u, 0u, 0u, 0u};
    for (uint32_t j = 0; j < Dt.m; ++j) {
      const uint32_t pos = Dt.pos[j];
      if (!Dt.tile[j]) idx0 |= (uint32_t)((cx.base >> pos) & 1ull) << j;
      else if (pos < 6) idx0 |= ((lane >> pos) & 1u) << j;
      else {
        slot_pos |= 1u << (pos - 6);
        slot_bit[0] = pos == 6 ? 1u << j : slot_bit[0];
        slot_bit[1] = pos == 7 ? 1u << j : slot_bit[1];
        slot_bit[2] = pos == 8 ? 1u << j : slot_bit[2];
        slot_bit[3] = pos == 9 ? 1u << j : slot_bit[3];
      }
    }
    (void)slot_pos;
    const uint64_t cbase = Dt.coef_off + idx0;
#pragma unroll
    for (int v = 0; v < KV; ++v) {
      QSIM_ITER_FENCE();
      const uint32_t sidx = ((v & 1) ? slot_bit[0] : 0u) | ((v & 2) ? slot_bit[1] : 0u) |
                            ((v & 4) ? slot_bit[2] : 0u) | ((v & 8) ? slot_bit[3] : 0u);
      amp[v] = cmul(amp[v], ldc(C, cbase + sidx));
    }
  }
}

// one DPP move of a double (both halves)
template <int CTRL, int RMASK>
__device__ __forceinline__ double dpp_f64(double v) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, RMASK, 0xf, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, RMASK, 0xf, false);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
// wave64 sum through DPP row shifts and row broadcasts; lane 63 ends with
// the total
__device__ __forceinline__ double wave_sum_f64(double x) {
  x += dpp_f64<0x111, 0xf>(x);  // row_shr:1
  x += dpp_f64<0x112, 0xf>(x);  // row_shr:2
  x += dpp_f64<0x114, 0xf>(x);  // row_shr:4
  x += dpp_f64<0x118, 0xf>(x);  // row_shr:8
  x += dpp_f64<0x142, 0xa>(x);  // row_bcast:15
  x += dpp_f64<0x143, 0xc>(x);  // row_bcast:31
  return x;
}

// ---------------------------------------------- points of the paired sweep
// Products of a gradient point on the paired kernel's lambda wave: element v
// of the lane pairs lambda row v with the parked psi element at row v ^ R,
// lane lane ^ (flip mask & 63).  t[v] = Re(conj(l) x) (im false) or
// Im(conj(l) x) (im true).  R is a template argument so that every LDS read
// has a constant offset from the lane's address.
template <int KV, int R>
__device__ __forceinline__ void pair_point_products(const double2* xr, const double2 (&amp)[KV], bool im,
                                                    double (&t)[KV]) {
  if (im) {
#pragma unroll
    for (int v = 0; v < KV; ++v) {
      const double2 x = xr[64 * (v ^ R)], l = amp[v];
      t[v] = fma(l.x, x.y, -(l.y * x.x));
    }
  } else {
#pragma unroll
    for (int v = 0; v < KV; ++v) {
      const double2 x = xr[64 * (v ^ R)], l = amp[v];
      t[v] = fma(l.x, x.x, l.y * x.y);
    }
  }
}
template <int KV>
__device__ __forceinline__ void pair_point_products_r(const double2* xr, const double2 (&amp)[KV], uint32_t r, bool im,
                                                      double (&t)[KV]) {
  switch (r) {
    case 0: pair_point_products<KV, 0>(xr, amp, im, t); break;
    case 1: pair_point_products<KV, 1>(xr, amp, im, t); break;
    case 2: pair_point_products<KV, 2>(xr, amp, im, t); break;
    case 3: pair_point_products<KV, 3>(xr, amp, im, t); break;
    case 4: pair_point_products<KV, 4>(xr, amp, im, t); break;
    case 5: pair_point_products<KV, 5>(xr, amp, im, t); break;
    case 6: pair_point_products<KV, 6>(xr, amp, im, t); break;
    case 7: pair_point_products<KV, 7>(xr, amp, im, t); break;
    default:
      if constexpr (KV == 16) {
        switch (r) {
          case 8: pair_point_products<KV, 8>(xr, amp, im, t); break;
          case 9: pair_point_products<KV, 9>(xr, amp, im, t); break;
          case 10: pair_point_products<KV, 10>(xr, amp, im, t); break;
          case 11: pair_point_products<KV, 11>(xr, amp, im, t); break;
          case 12: pair_point_products<KV, 12>(xr, amp, im, t); break;
          case 13: pair_point_products<KV, 13>(xr, amp, im, t); break;
          case 14: pair_point_products<KV, 14>(xr, amp, im, t); break;
          default: pair_point_products<KV, 15>(xr, amp, im, t); break;
        }
      }
      break;
  }
}
// sum over v of (-1)^popc(v & zr) t[v]: one level per row bit, the rows with
// the bit set taken with the sign of zr's bit (a Walsh coefficient of t)
template <int KV>
__device__ __forceinline__ double walsh_fold(const double (&t)[KV], uint32_t zr) {
  double f[KV];
#pragma unroll
  for (int v = 0; v < KV; ++v) f[v] = t[v];
#pragma unroll
  for (int h = 1; h < KV; h <<= 1) {
    const double sg = (zr & (uint32_t)h) ? -1.0 : 1.0;
#pragma unroll
    for (int v = 0; v < KV; v += 2 * h) f[v] = fma(sg, f[v + h], f[v]);
  }
  return f[0];
}
// In place: t[z] <- sum_v (-1)^popc(v & z) t[v] (the Walsh-Hadamard
// transform over the row bits; entry z is walsh_fold(t, z))
template <int KV>
__device__ __forceinline__ void walsh_transform(double (&t)[KV]) {
#pragma unroll
  for (int h = 1; h < KV; h <<= 1) {
#pragma unroll
    for (int v = 0; v < KV; ++v) {
      if (v & h) continue;
      const double x = t[v], y = t[v + h];
      t[v] = x + y;
      t[v + h] = x - y;
    }
  }
}

// ---------------------------------------------- points of lambda-only passes
// Rows of one gradient point Re<lambda| i P |psi> on a tile (lambda in lam,
// psi in ps at the same positions): P flips tile rows by FR and lanes by fl,
// with the Z/Y sign of the partner index and Re or Im by kr, as the paired
// kernel forms it.
template <int FR, int KV>
__device__ __forceinline__ void point_rows(const double2 (&lam)[KV], const double2 (&ps)[KV], uint32_t fl, uint32_t fx,
                                           uint32_t zt, uint32_t kr, uint32_t lsgn, double& acc0, double& acc1) {
  if constexpr (FR >= KV) {
    return;  // no such row flip in this tile shape
  } else {
    const bool odd = (kr & 1u) != 0u;
#pragma unroll
    for (int v = 0; v < KV; ++v) {
      double2 x = ps[v ^ FR];
      if (fl) x = shfl_xor2(x, (int)fl);
      const double2 l = lam[v];
      const uint32_t sg = lsgn ^ ((uint32_t)(__popc((((uint32_t)v << 6) ^ (fx & ~63u)) & zt) & 1) << 31);
      double t = odd ? fma(l.x, x.y, -(l.y * x.x)) : fma(l.x, x.x, l.y * x.y);
      t = __longlong_as_double(__double_as_longlong(t) ^ ((long long)sg << 32));
      if (v & 1) acc1 += t;
      else acc0 += t;
    }
  }
}
// The tile sum of a gradient point (the same value in every lane).
template <int KV>
__device__ __forceinline__ double point_tile_sum(const QC Op& op, const double2 (&lam)[KV], const double2 (&ps)[KV],
                                                 uint32_t lane, uint64_t base) {
  const uint32_t fx = (uint32_t)op.gfx, zt = (uint32_t)op.gzt;
  const uint32_t kr = (1u + op.nY) & 3u;  // i * i^nY = i^kr
  const bool flip_all = (__popcll(base & op.gzb) & 1) ^ (kr == 1u || kr == 2u);
  const uint32_t fl = fx & 63u, fr = fx >> 6;
  const uint32_t lsgn = (uint32_t)(__popc((lane ^ fx) & zt & 63u) & 1) << 31;
  double acc0 = 0.0, acc1 = 0.0;
  switch (fr) {
    case 0: point_rows<0>(lam, ps, fl, fx, zt, kr, lsgn, acc0, acc1); break;
    case 1: point_rows<1>(lam, ps, fl, fx, zt, kr, lsgn, acc0, acc1); break;
    case 2: point_rows<2>(lam, ps, fl, fx, zt, kr, lsgn, acc0, acc1); break;
    case 3: point_rows<3>(lam, ps, fl, fx, zt, kr, lsgn, acc0, acc1); break;
    case 4: point_rows<4>(lam, ps, fl, fx, zt, kr, lsgn, acc0, acc1); break;
    case 5: point_rows<5>(lam, ps, fl, fx, zt, kr, lsgn, acc0, acc1); break;
    case 6: point_rows<6>(lam, ps, fl, fx, zt, kr, lsgn, acc0, acc1); break;
    case 7: point_rows<7>(lam, ps, fl, fx, zt, kr, lsgn, acc0, acc1); break;
    case 8: point_rows<8>(lam, ps, fl, fx, zt, kr, lsgn, acc0, acc1); break;
    case 9: point_rows<9>(lam, ps, fl, fx, zt, kr, lsgn, acc0, acc1); break;
    case 10: point_rows<10>(lam, ps, fl, fx, zt, kr, lsgn, acc0, acc1); break;
    case 11: point_rows<11>(lam, ps, fl, fx, zt, kr, lsgn, acc0, acc1); break;
    case 12: point_rows<12>(lam, ps, fl, fx, zt, kr, lsgn, acc0, acc1); break;
    case 13: point_rows<13>(lam, ps, fl, fx, zt, kr, lsgn, acc0, acc1); break;
    case 14: point_rows<14>(lam, ps, fl, fx, zt, kr, lsgn, acc0, acc1); break;
    default: point_rows<15>(lam, ps, fl, fx, zt, kr, lsgn, acc0, acc1); break;
  }
  const double tot = __shfl(wave_sum_f64(acc0 + acc1), 63, 64);
  return flip_all ? -tot : tot;
}

// ------------------------------------------------------------- the kernel
// KV complex rows per lane: a wave tile of 64 KV amplitudes (lanes on local
// qubits 0..5, log2 KV register qubits).  KV = 16 is the default tile; KV = 8
// halves the registers of a wave so that three waves fit a SIMD, for stages
// whose non-diagonal gates need only three register qubits.  In chunk layout
// 32/KV lanes share one 32-amplitude codec chunk: lane L holds amplitudes
// KV g .. KV g + KV - 1 (g = L mod 32/KV) of chunk L / (32/KV).
// PT: a lambda-only reverse pass (RAW->RAW on lambda, 16 rows) whose op list
// may start with gradient points, evaluated against the stored psi.
template <int IN, int OUT, int KV, bool PT = false>
__global__ __launch_bounds__(kThreads, KV == 16 ? 2 : 3) void wave_kernel(const PassArgs* __restrict__ pa, PassDyn dy,
                                                                          uint64_t ntiles) {
  static_assert(!PT || (IN == IN_RAW && OUT == OUT_RAW && KV == 16), "points run on 16-row RAW passes");
  constexpr int RB = KV == 16 ? 4 : 3;            // register qubits
  constexpr uint32_t TB = 6 + RB;                 // tile bits
  constexpr uint32_t LPC = 32 / KV;               // lanes per chunk
  constexpr uint32_t XS = x_slots<KV>();
  // per wave: the layout transposes (XS complex128) and, reused after the
  // encoder's transpose, the packed payload of the tile (at most 128 KV
  // words) plus slack for the zero-valued spill OR past the last field and
  // one spare word per lane
  static_assert(2 * XS >= 128 * KV + 8 + 64, "staging and the lanes' spare words must fit the transpose buffer");
  __shared__ __align__(16) uint64_t wbuf[kThreads / 64][2 * XS];
  const QC PassArgs& a = *(const QC PassArgs*)(pa);
  const QC Op* ops = (const QC Op*)(a.ops);
  const QC PhaseTerm* terms = (const QC PhaseTerm*)(a.terms);
  const QC DiagTable* diags = (const QC DiagTable*)(a.diags);
  const QC double* coefs = (const QC double*)(a.coefs);
  const QG uint32_t* bmap = (const QG uint32_t*)(dy.bmap);
  const QG uint8_t* null_in = (const QG uint8_t*)(dy.null_in);
  const QG uint64_t* desc_in = (const QG uint64_t*)(dy.desc_in);
  const QG uint64_t* arena_in = (const QG uint64_t*)(dy.arena_in);
  QG uint64_t* desc_out = (QG uint64_t*)(dy.desc_out);
  QG uint64_t* arena_out = (QG uint64_t*)(dy.arena_out);
  QG uint8_t* nonzero = (QG uint8_t*)(dy.nonzero);
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  // The forms that decode run a persistent grid (wave w of group b takes
  // tiles tile0 + 4 (b + k gridDim) + w) and prefetch the next tile's chunk
  // metadata.  RAW->ENC takes one tile per wave: its persistent form spills
  // 16 B per lane and ran 4% slower (QFT-32's 2^32-amplitude last pass,
  // 22.5 ms against 21.7 ms).
  constexpr bool kPersistent = IN == IN_DEC;
  // RAW->RAW passes on 1024-amplitude tiles also run persistently and
  // pipeline their input: the rows of a wave's next tile are copied into the
  // wave's LDS buffer (unused by this form) by global-to-LDS loads while the
  // current tile's gates run.
  constexpr bool kPipeRaw = IN == IN_RAW && OUT == OUT_RAW && KV == 16;
  static_assert(!kPipeRaw || 2 * XS * sizeof(uint64_t) >= 64 * KV * sizeof(double2), "staging rows must fit");
  auto stage_rows = [&](const uint64_t wt) {
    const uint64_t slot = wt >> (a.L - TB);
    const TileMapConst tm{TB, 6u, (uint32_t)RB, a.E};
    const uint64_t base = tm.base(wt & ((1ull << (a.L - TB)) - 1ull));
    const QG double2* src = (const QG double2*)(dy.raw) + (slot << a.L) + base + lane;
    __attribute__((address_space(3))) double2* dst = (__attribute__((address_space(3))) double2*)(&wbuf[wave][0]);
#pragma unroll
    for (int v = 0; v < KV; ++v) {
      const uint64_t ro = ((v & 1) ? 1ull << a.E[0] : 0ull) | ((v & 2) ? 1ull << a.E[1] : 0ull) |
                          ((v & 4) ? 1ull << a.E[2] : 0ull) | ((v & 8) ? 1ull << a.E[3] : 0ull);
      __builtin_amdgcn_global_load_lds((const void*)(src + ro), (__attribute__((address_space(3))) void*)(dst + 64 * v),
                                       16, 0, 0);
    }
  };
  // Wave-owned sub-arenas (persistent DEC->ENC on the double store, when
  // dy.wave_subs is set): wave w of the grid packs the payloads of all its
  // tiles back to back into sub-arena w, so the allocation is a running sum
  // kept in a scalar register instead of a device-scope atomic per tile.
  const bool own = kPersistent && OUT == OUT_ENC && dy.wave_subs != 0u;
  const uint32_t gwave = (uint32_t)blockIdx.x * (kThreads / 64) + (uint32_t)wave;
  uint64_t wcur = 0;
  // a lambda-only pass whose op list starts with gradient points
  const bool pts = PT && a.op_begin < a.op_end && ops[a.op_begin].kind == OP_GRAD;
  // this lane's slot offset of an ENC store's tile permutation
  const uint32_t tperm_sl = OUT == OUT_ENC && a.tperm ? (uint32_t)a.tperm_lane[lane] : 0u;
  // Chunk metadata of tile wt (chunk layout): the LPC lanes of chunk c all
  // describe it; its first amplitude is tile position 32c.  The block-map
  // load (meta_bpos) and the null-flag and descriptor loads that depend on it
  // (meta_rest) are separate steps, so that the persistent DEC->ENC loop can
  // issue a tile's block-map load one tile early and its other loads during
  // the previous tile's encode.
  struct TileMeta {
    uint64_t desc;
    uint32_t bid, ci, nul;
  };
  auto meta_loc = [&](const uint64_t wt) -> uint64_t {
    const TileMapConst tm{TB, 6u, (uint32_t)RB, a.E};
    return tm.base(wt & ((1ull << (a.L - TB)) - 1ull)) | tm.off(32u * ((uint32_t)lane / LPC));
  };
  auto meta_bpos = [&](const uint64_t wt) -> uint64_t {
    return ((wt >> (a.L - TB)) << a.khi) + (meta_loc(wt) >> a.B);
  };
  auto meta_rest = [&](const uint64_t wt, const uint32_t bid) -> TileMeta {
    const uint64_t slot = wt >> (a.L - TB), loc = meta_loc(wt);
    const uint64_t bpos = (slot << a.khi) + (loc >> a.B);
    TileMeta m{0ull, bid, (uint32_t)((loc & ((1ull << a.B) - 1ull)) >> kChunkAmpsLog2), 1u};
    m.nul = null_in[bid];
    if constexpr (IN == IN_DEC) {
      // the descriptor table spans every block, so the descriptor load need
      // not wait for the null flag (a null block's stale entry is discarded)
      const uint64_t d = desc_in[(dy.desc_staged ? bpos : (uint64_t)bid) * a.cpb + m.ci];
      m.desc = m.nul ? 0ull : d;
    }
    if constexpr (OUT == OUT_ENC) {
      // the store's chunk permutation: from here on the metadata is that of
      // the chunk's output position
      if (a.nswap) {
        uint64_t lo = loc;
        for (uint32_t i = 0; i < a.nswap; ++i) {
          const uint32_t p0 = a.swp[2 * i], p1 = a.swp[2 * i + 1];
          const uint64_t d = ((lo >> p0) ^ (lo >> p1)) & 1ull;
          lo ^= (d << p0) | (d << p1);
        }
        m.bid = bmap[(slot << a.khi) + (lo >> a.B)];
        m.ci = (uint32_t)((lo & ((1ull << a.B) - 1ull)) >> kChunkAmpsLog2);
        m.nul = null_in[m.bid];
      }
    }
    return m;
  };
  auto meta_of = [&](const uint64_t wt) -> TileMeta {
    if constexpr (IN == IN_DEC || OUT == OUT_ENC) return meta_rest(wt, bmap[meta_bpos(wt)]);
    else return TileMeta{0ull, 0u, 0u, 1u};
  };
  // pre(): for RAW->RAW on 1024-amplitude tiles, the next tile's row loads
  // (called after this tile's rows are in registers); for the persistent
  // DEC->ENC loop, the next tile's metadata loads (called after the ops)
  auto tile_body = [&](const uint64_t wt, const TileMeta m, auto&& pre) {
  const uint32_t L = a.L;
  const uint64_t slot = wt >> (L - TB);
  const uint64_t tau = wt & ((1ull << (L - TB)) - 1ull);
  const TileMapConst tm{TB, 6u, (uint32_t)RB, a.E};
  const uint64_t base = tm.base(tau);
  const uint64_t e0 = 1ull << a.E[0], e1 = 1ull << a.E[1], e2 = 1ull << a.E[2];
  const uint64_t e3 = RB > 3 ? 1ull << a.E[3] : 0ull;
  auto roff = [&](int v) -> uint64_t {
    return ((v & 1) ? e0 : 0ull) | ((v & 2) ? e1 : 0ull) | ((v & 4) ? e2 : 0ull) | ((v & 8) ? e3 : 0ull);
  };
  const WaveCtx cx{lane, base, tau, a.ptab};
  double2 amp[KV];
  [[maybe_unused]] double2 ps[PT ? KV : 1];
  double2* xb = reinterpret_cast<double2*>(wbuf[wave]);
  const uint32_t g = (uint32_t)lane & (LPC - 1u);  // part of the chunk (chunk layout)
  const uint32_t d0 = 2u * KV * g;                 // first double of this lane's part

  const uint64_t mdesc = m.desc;
  const uint32_t mbid = m.bid, mci = m.ci, mnull = m.nul;

  // ---- load
  if constexpr (kPipeRaw) {
    // this tile's rows were staged by stage_rows; every earlier vector memory
    // operation of the wave (those loads, the previous tile's stores) done
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    const double2* pf = reinterpret_cast<const double2*>(wbuf[wave]);
#pragma unroll
    for (int v = 0; v < KV; ++v) amp[v] = pf[64 * v + lane];
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    if constexpr (PT) {
      // psi rows of the points, issued before the next tile's rows so that
      // waiting for them does not wait for those
      if (pts) {
        const QG double* sp = (const QG double*)(dy.psi) + 2 * ((slot << L) + base + lane);
#pragma unroll
        for (int v = 0; v < KV; ++v) ps[v] = make_double2(sp[2 * roff(v)], sp[2 * roff(v) + 1]);
      }
    }
    pre();  // the next tile's rows into the same buffer
    __builtin_amdgcn_sched_barrier(0);
  } else if constexpr (IN == IN_RAW) {
    const QG double* src = (const QG double*)(dy.raw) + 2 * ((slot << L) + base + lane);
#pragma unroll
    for (int v = 0; v < KV; ++v) amp[v] = make_double2(src[2 * roff(v)], src[2 * roff(v) + 1]);
  } else {
    // this lane's codes are doubles d0 .. d0 + 2 KV - 1 of its chunk
    const uint32_t w = desc_w(mdesc);
    if (ballot64(w > 32u) == 0) {
      // every code of the tile has at most 32 bits: all payload loads are
      // issued before the first decode (one memory round trip per tile)
      const int e = desc_e(mdesc);
      const QG uint32_t* P32 = (const QG uint32_t*)(arena_in + desc_off(mdesc));
      const uint32_t m = (uint32_t)((1ull << w) - 1ull);
      uint32_t dw[KV][3];
#pragma unroll
      for (int q = 0; q < KV; ++q) {
        const uint32_t bit = (d0 + 2u * q) * w;
        const QG uint32_t* P = P32 + (bit >> 5);
        dw[q][0] = P[0];
        dw[q][1] = P[1];
        dw[q][2] = P[2];
      }
#pragma unroll
      for (int q = 0; q < KV; ++q) {
        const uint32_t sh = ((d0 + 2u * q) * w) & 31u;
        const uint32_t lo = __builtin_amdgcn_alignbit(dw[q][1], dw[q][0], sh);
        const uint32_t hi = __builtin_amdgcn_alignbit(dw[q][2], dw[q][1], sh);
        const uint64_t x = ((uint64_t)hi << 32) | lo;
        const uint32_t u0 = lo & m, u1 = (uint32_t)(x >> w) & m;
        const int k0 = (int)(u0 >> 1) ^ -(int)(u0 & 1u), k1 = (int)(u1 >> 1) ^ -(int)(u1 & 1u);
        amp[q] = make_double2(ldexp((double)k0, e), ldexp((double)k1, e));
      }
    } else {
#pragma unroll
      for (int q = 0; q < KV; ++q)
        amp[q] = make_double2(decode_at(mdesc, arena_in, d0 + 2u * q), decode_at(mdesc, arena_in, d0 + 2u * q + 1u));
    }
    to_wave_layout(amp, xb, (uint32_t)lane);
  }

  // ---- ops
  uint32_t o0 = a.op_begin;
  if constexpr (PT) {
    // the points at the start of a lambda-only pass: amp holds lambda, ps psi
    if (pts) {
      for (; o0 < a.op_end && ops[o0].kind == OP_GRAD; ++o0) {
        const QC Op& op = ops[o0];
        const double tot = point_tile_sum<KV>(op, amp, ps, (uint32_t)lane, base);
        if (lane == 0) {
          dy.partials[((uint64_t)op.gidx * dy.part_tiles + wt) * 2] = tot;
          dy.partials[((uint64_t)op.gidx * dy.part_tiles + wt) * 2 + 1] = 0.0;
        }
      }
    }
  }
  for (uint32_t o = o0; o < a.op_end; ++o) {
    const QC Op& op = ops[o];
    if (op.kind == OP_DENSE) wave_dense(op, amp, cx, coefs);
    else wave_elem(op, amp, cx, terms, diags, coefs);
  }
  if constexpr (kPersistent) pre();

  // ---- store
  if constexpr (OUT == OUT_RAW) {
    QG double* dst = (QG double*)(dy.raw_out ? dy.raw_out : dy.raw) + 2 * ((slot << L) + base + lane);
#pragma unroll
    for (int v = 0; v < KV; ++v) {
      dst[2 * roff(v)] = amp[v].x;
      dst[2 * roff(v) + 1] = amp[v].y;
    }
  } else {
    if (a.tperm) to_chunk_layout_perm(amp, xb, (uint32_t)lane, tperm_sl, a.tperm_row);
    else to_chunk_layout(amp, xb, (uint32_t)lane);
    // zero check of blocks that were null at stage start
    if (ballot64(mnull != 0u) != 0) {
      bool nz = false;
#pragma unroll
      for (int q = 0; q < KV; ++q)
        nz = nz || (a.zero_mode == ZERO_EXACT ? (amp[q].x != 0.0 || amp[q].y != 0.0)
                                              : (fabs(amp[q].x) > a.zero_thr || fabs(amp[q].y) > a.zero_thr));
      if (mnull && nz) nonzero[mbid] = 1;
    }
    const CodecParams cp{a.eps_rel, a.eps_abs, a.guard, a.e_abs, a.p_abs};
    const uint32_t plan = lane_quantise(amp, cp, a.ovf_hi);
    const uint32_t w = plan & 0xffu;
    // payload offsets: chunks in order, each w words
    const uint32_t incl = wave_incl_scan(g == 0 ? w : 0u);
    const uint32_t loff = incl - w;
    const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    // one allocation per wave tile; its latency is hidden behind the packing.
    // The region store routes a group's tiles to its output region's
    // sub-arenas (pairs of bounds: sub-arena s at sub_base[2s], [2s + 1]).
    const uint32_t sub = dy.slot_region ? 2u * (dy.slot_region[slot] * (dy.sub_mask + 1u) + region_piece(wt, dy.sub_mask))
                                        : (uint32_t)(wt & dy.sub_mask);
    unsigned long long aoff = 0;
    if (own) {
      aoff = wcur;
      wcur += tot;
    } else if (lane == 0 && tot > 0) {
      aoff = atomicAdd(&dy.cursors[sub * 16], (unsigned long long)tot);
    }
    uint64_t* st = wbuf[wave];
    for (uint32_t q = lane; q < tot; q += 64) st[q] = 0ull;
    __builtin_amdgcn_wave_barrier();
    uint64_t* sc = st + loff;
    if (ballot64(w > 32u) == 0) {
      // codes of at most 32 bits: pieces of as many codes as the tile's
      // widest chunk allows in 64 bits; each lane's spare word (past the
      // staging area) takes the ORs that would add nothing
      uint64_t* dummy = st + 128 * KV + 8 + lane;
      if (ballot64(w > 16u)) pack_pieces<KV, 2>(sc, dummy, amp, d0, w);
      else if (ballot64(w > 8u)) pack_pieces<KV, 4>(sc, dummy, amp, d0, w);
      else if (ballot64(w > 4u)) pack_pieces<KV, 8>(sc, dummy, amp, d0, w);
      else pack_pieces<KV, 16>(sc, dummy, amp, d0, w);
    } else {
#pragma unroll
      for (int q = 0; q < KV; ++q) {
        const uint32_t di = d0 + 2u * q;
        stage_or(sc, di * w, (uint64_t)__double_as_longlong(amp[q].x));
        stage_or(sc, (di + 1u) * w, (uint64_t)__double_as_longlong(amp[q].y));
      }
    }
    __builtin_amdgcn_wave_barrier();
    if (tot > 0) {
      const uint64_t off = readlane64(aoff, 0);
      const uint32_t asub = own ? gwave : sub;
      const uint64_t sb0 = dy.sub_base[asub], sb1 = dy.sub_base[asub + 1];
      if (off + tot > sb1 - sb0) {
        if (lane == 0) atomicOr(dy.overflow, 1u);
      } else {
        const uint64_t tbase = off + sb0;
        for (uint32_t q = lane; q < tot; q += 64) arena_out[tbase + q] = st[q];
        if (g == 0) desc_out[(uint64_t)mbid * a.cpb + mci] = desc_pack(tbase + loff, w, (int)(int16_t)(uint16_t)(plan >> 8));
      }
    } else if (g == 0) {
      desc_out[(uint64_t)mbid * a.cpb + mci] = desc_pack(0, 0, 0);
    }
  }
  };
  const uint64_t wt0 = dy.tile0 + (uint64_t)blockIdx.x * (kThreads / 64) + wave;
  if constexpr (kPipeRaw) {
    const uint64_t stride = (uint64_t)gridDim.x * (kThreads / 64);
    if (wt0 < ntiles) stage_rows(wt0);
    for (uint64_t wt = wt0; wt < ntiles; wt += stride)
      tile_body(wt, meta_of(wt), [&] {
        if (wt + stride < ntiles) stage_rows(wt + stride);
      });
  } else if constexpr (kPersistent) {
    const uint64_t stride = (uint64_t)gridDim.x * (kThreads / 64);
    if (wt0 < ntiles) {
      TileMeta m = meta_of(wt0);
      for (uint64_t wt = wt0; wt < ntiles; wt += stride) {
        const uint64_t wn = wt + stride;
        const bool more = wn < ntiles;  // uniform over the wave
        uint32_t nbid = 0;
        if (more) nbid = bmap[meta_bpos(wn)];
        TileMeta mn{0ull, 0u, 0u, 1u};
        tile_body(wt, m, [&] {
          if (more) mn = meta_rest(wn, nbid);
        });
        m = mn;
      }
    }
    // the sub-arena's demand (all tiles counted, also past an overflow)
    if (own && lane == 0) dy.cursors[(uint64_t)gwave * 16] = wcur;
  } else {
    if (wt0 < ntiles) tile_body(wt0, meta_of(wt0), [] {});
  }
}

// --------------------------------------------------- adjoint reverse sweep
// Re(i^kr * conj(l) * x * sign) for one element of a gradient point
__device__ __forceinline__ double grad_term(double2 l, double2 x, bool neg, int kr) {
  double re = l.x * x.x + l.y * x.y, im = l.x * x.y - l.y * x.x;
  re = neg ? -re : re;
  im = neg ? -im : im;
  return kr == 0 ? re : kr == 1 ? -im : kr == 2 ? -re : im;
}
__device__ __forceinline__ void dd_acc(double& s, double& c, double v) {
  const double t = s + v, bp = t - s;
  c += (s - (t - bp)) + (v - bp);
  s = t;
}
// Pauli partner on register qubits only: psi row v pairs with row v ^ M;
// lambda lives in rows kV/2 .. kV-1
template <int M>
__device__ __forceinline__ void grad_rows(const double2 (&amp)[kV], uint32_t zt, bool bneg, int kr, int lane,
                                          double& s, double& c) {
  constexpr int R = kV / 2;
#pragma unroll
  for (int v = 0; v < R; ++v) {
    const uint32_t q = ((uint32_t)(v ^ M) << 6) | (uint32_t)lane;
    dd_acc(s, c, grad_term(amp[R + v], amp[v ^ M], bneg ^ (bool)(__popc(q & zt) & 1), kr));
  }
}
// Single-wave form of the adjoint pass (ParametricQuantumCircuit::
// backprop_inner_product, src/vqcsim/parametric_circuit.cpp:180-279) on
// 512-amplitude tiles: lanes on qubits 0..5 and three register qubits
// E[0..2].  Register rows 0..7 hold psi and rows 8..15 the same positions of
// lambda, so the fourth row bit selects the state and no gate ever acts on it:
// every dense / elementwise op (the inverted gates) runs once over all 16 rows
// with the forward kernel's code and register footprint.  An OP_GRAD op forms
// Re sum conj(lambda) (i P psi) over the tile from the Pauli partners (row
// swaps, lane shuffles, or the wave's LDS buffer for mixed flips) and writes it
// to partials[(gidx * ntiles + tile) * 2].
// One 512-amplitude tile wt of a single-wave adjoint pass: psi rows 0..7,
// lambda rows 8..15; sink(gidx, tot) receives every gradient point's tile
// sum (the same value in every lane).
template <class Sink>
__device__ __forceinline__ void adjoint_tile(const QC PassArgs& a, uint64_t wt, double2* __restrict__ psi,
                                             double2* __restrict__ lam, double2* xb, int lane, Sink&& sink) {
  constexpr int R = kV / 2;  // rows per state
  const QC Op* ops = (const QC Op*)(a.ops);
  const QC PhaseTerm* terms = (const QC PhaseTerm*)(a.terms);
  const QC DiagTable* diags = (const QC DiagTable*)(a.diags);
  const QC double* coefs = (const QC double*)(a.coefs);
  const TileMapConst tm{9u, 6u, 3u, a.E};
  const uint64_t base = tm.base(wt);
  const uint64_t e0 = 1ull << a.E[0], e1 = 1ull << a.E[1], e2 = 1ull << a.E[2];
  auto roff = [&](int v) -> uint64_t { return ((v & 1) ? e0 : 0ull) | ((v & 2) ? e1 : 0ull) | ((v & 4) ? e2 : 0ull); };
  const WaveCtx cx{lane, base, wt, a.ptab};
  double2 amp[kV];
  {
    const QG double* sp = (const QG double*)(psi) + 2 * (base + lane);
    const QG double* sl = (const QG double*)(lam) + 2 * (base + lane);
#pragma unroll
    for (int v = 0; v < R; ++v) {
      amp[v] = make_double2(sp[2 * roff(v)], sp[2 * roff(v) + 1]);
      amp[R + v] = make_double2(sl[2 * roff(v)], sl[2 * roff(v) + 1]);
    }
  }
  for (uint32_t o = a.op_begin; o < a.op_end; ++o) {
    const QC Op& op = ops[o];
    if (op.kind == OP_DENSE) {
      wave_dense(op, amp, cx, coefs);
    } else if (op.kind == OP_ELEM) {
      wave_elem(op, amp, cx, terms, diags, coefs);
    } else {
      const uint32_t fx = (uint32_t)op.gfx, zt = (uint32_t)op.gzt;
      const int kr = (int)((1u + op.nY) & 3u);  // i * i^nY = i^kr
      const bool bneg = __popcll(base & op.gzb) & 1;
      double s = 0.0, c = 0.0;
      if (fx == 0u) {
#pragma unroll
        for (int v = 0; v < R; ++v) {
          const uint32_t q = ((uint32_t)v << 6) | (uint32_t)lane;
          dd_acc(s, c, grad_term(amp[R + v], amp[v], bneg ^ (bool)(__popc(q & zt) & 1), kr));
        }
      } else if ((fx & 63u) == 0u) {
        switch (fx >> 6) {
          case 1: grad_rows<1>(amp, zt, bneg, kr, lane, s, c); break;
          case 2: grad_rows<2>(amp, zt, bneg, kr, lane, s, c); break;
          case 3: grad_rows<3>(amp, zt, bneg, kr, lane, s, c); break;
          case 4: grad_rows<4>(amp, zt, bneg, kr, lane, s, c); break;
          case 5: grad_rows<5>(amp, zt, bneg, kr, lane, s, c); break;
          case 6: grad_rows<6>(amp, zt, bneg, kr, lane, s, c); break;
          default: grad_rows<7>(amp, zt, bneg, kr, lane, s, c); break;
        }
      } else if (fx < 64u) {
#pragma unroll
        for (int v = 0; v < R; ++v) {
          QSIM_ITER_FENCE();
          const uint32_t q = (((uint32_t)v << 6) | (uint32_t)lane) ^ fx;
          dd_acc(s, c, grad_term(amp[R + v], shfl_xor2(amp[v], (int)fx), bneg ^ (bool)(__popc(q & zt) & 1), kr));
        }
      } else {
        // flips on lanes and register qubits: psi through the wave's LDS
        // buffer (position 64v + l at slot 68v + l + l/16)
#pragma unroll
        for (int v = 0; v < R; ++v) xb[68 * v + lane + (lane >> 4)] = amp[v];
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int v = 0; v < R; ++v) {
          const uint32_t q = (((uint32_t)v << 6) | (uint32_t)lane) ^ fx;
          dd_acc(s, c, grad_term(amp[R + v], xb[q + (q >> 4)], bneg ^ (bool)(__popc(q & zt) & 1), kr));
        }
        __builtin_amdgcn_wave_barrier();
      }
      // tile sum: the lanes' compensated sums added in a DPP tree
      sink(op.gidx, __shfl(wave_sum_f64(s + c), 63, 64));
    }
  }
  {
    QG double* dp = (QG double*)(psi) + 2 * (base + lane);
    QG double* dl = (QG double*)(lam) + 2 * (base + lane);
#pragma unroll
    for (int v = 0; v < R; ++v) {
      dp[2 * roff(v)] = amp[v].x;
      dp[2 * roff(v) + 1] = amp[v].y;
      dl[2 * roff(v)] = amp[R + v].x;
      dl[2 * roff(v) + 1] = amp[R + v].y;
    }
  }
}

__global__ __launch_bounds__(kThreads, 2) void wave_adjoint_kernel(const PassArgs* __restrict__ pa,
                                                                   double2* __restrict__ psi,
                                                                   double2* __restrict__ lam,
                                                                   double* __restrict__ partials, uint64_t tile0,
                                                                   uint64_t ntiles) {
  __shared__ __align__(16) double2 xbuf[kThreads / 64][4 * 128 + 32];
  const QC PassArgs& a = *(const QC PassArgs*)(pa);
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint64_t wt = tile0 + (uint64_t)blockIdx.x * (kThreads / 64) + wave;
  if (wt >= ntiles) return;
  adjoint_tile(a, wt, psi, lam, xbuf[wave], lane, [&](uint32_t g, double tot) {
    if (lane == 0) {
      partials[((uint64_t)g * ntiles + wt) * 2] = tot;
      partials[((uint64_t)g * ntiles + wt) * 2 + 1] = 0.0;
    }
  });
}

// One 1024-amplitude tile wt of a RAW->RAW wave pass (rows loaded and stored
// directly), for the persistent batch kernel.
__device__ __forceinline__ void raw_tile16(const QC PassArgs& a, uint64_t wt, double2* __restrict__ raw, int lane) {
  const QC Op* ops = (const QC Op*)(a.ops);
  const QC PhaseTerm* terms = (const QC PhaseTerm*)(a.terms);
  const QC DiagTable* diags = (const QC DiagTable*)(a.diags);
  const QC double* coefs = (const QC double*)(a.coefs);
  const TileMapConst tm{10u, 6u, 4u, a.E};
  const uint64_t base = tm.base(wt);
  const uint64_t e0 = 1ull << a.E[0], e1 = 1ull << a.E[1], e2 = 1ull << a.E[2], e3 = 1ull << a.E[3];
  auto roff = [&](int v) -> uint64_t {
    return ((v & 1) ? e0 : 0ull) | ((v & 2) ? e1 : 0ull) | ((v & 4) ? e2 : 0ull) | ((v & 8) ? e3 : 0ull);
  };
  const WaveCtx cx{lane, base, wt, a.ptab};
  double2 amp[16];
  QG double* sp = (QG double*)(raw) + 2 * (base + lane);
#pragma unroll
  for (int v = 0; v < 16; ++v) amp[v] = make_double2(sp[2 * roff(v)], sp[2 * roff(v) + 1]);
  for (uint32_t o = a.op_begin; o < a.op_end; ++o) {
    const QC Op& op = ops[o];
    if (op.kind == OP_DENSE) wave_dense(op, amp, cx, coefs);
    else wave_elem(op, amp, cx, terms, diags, coefs);
  }
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    sp[2 * roff(v)] = amp[v].x;
    sp[2 * roff(v) + 1] = amp[v].y;
  }
}

// Persistent batch kernel (the batched adjoint of qsim_batch_run, per
// instance ParametricQuantumCircuit::backprop semantics,
// src/vqcsim/parametric_circuit.cpp:268-308): each work-group takes whole
// instances from *next until none is left and runs one start to end, the
// passes of an instance separated by work-group barriers (its tiles are all
// in this work-group, so no other synchronisation is needed).  Every sum of
// an instance is formed in a fixed order inside its work-group: results do
// not depend on which work-group ran it.
__global__ __launch_bounds__(kThreads, 2) void batch_persistent_kernel(const BatchJob* __restrict__ jobs,
                                                                       uint32_t njobs, unsigned* __restrict__ next) {
  __shared__ __align__(16) double2 xbuf[kThreads / 64][4 * 128 + 32];
  __shared__ double pts[kBatchMaxPoints][kThreads / 64];
  __shared__ double red[kThreads / 64];
  __shared__ uint32_t sj;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  constexpr uint32_t W = kThreads / 64;
  for (uint32_t q = threadIdx.x; q < kBatchMaxPoints * W; q += kThreads) pts[q / W][q % W] = 0.0;
  for (;;) {
    if (threadIdx.x == 0) sj = atomicAdd(next, 1u);
    __syncthreads();
    const uint32_t j = (uint32_t)__builtin_amdgcn_readfirstlane((int)sj);
    __syncthreads();
    if (j >= njobs) return;  // uniform over the work-group
    const QC BatchJob& J = *(const QC BatchJob*)(jobs + j);
    const uint64_t dim = 1ull << J.n;
    double2* psi = J.psi;
    double2* lam = J.lam;
    for (uint64_t x = threadIdx.x; x < dim; x += kThreads) psi[x] = make_double2(x == 0 ? 1.0 : 0.0, 0.0);
    __syncthreads();
    for (uint32_t p = 0; p < J.nfwd; ++p) {
      const QC PassArgs& a = *(const QC PassArgs*)(J.fwd + p);
      for (uint64_t wt = (uint64_t)wave; wt < (dim >> 10); wt += W) raw_tile16(a, wt, psi, lane);
      __syncthreads();
    }
    // lambda = H psi (Observable::apply_to_state) and <psi|lambda>
    {
      const QC PauliGroup* groups = (const QC PauliGroup*)(J.groups);
      const QC PauliTermZ* terms = (const QC PauliTermZ*)(J.terms);
      double s = 0.0, c = 0.0;
      for (uint64_t x = threadIdx.x; x < dim; x += kThreads) {
        double re = 0.0, im = 0.0;
        for (uint32_t g = 0; g < J.ng; ++g) {
          const QC PauliGroup& G = groups[g];
          const uint64_t y = x ^ G.flip;
          double fr = 0.0, fi = 0.0;
          for (uint32_t t = G.t0; t < G.t0 + G.nt; ++t) {
            const bool neg = __popcll(y & terms[t].z) & 1;
            fr += neg ? -terms[t].cr : terms[t].cr;
            fi += neg ? -terms[t].ci : terms[t].ci;
          }
          const double2 v = psi[y];
          re += fr * v.x - fi * v.y;
          im += fr * v.y + fi * v.x;
        }
        lam[x] = make_double2(re, im);
        const double2 ax = psi[x];
        dd_acc(s, c, ax.x * re + ax.y * im);
      }
      const double tot = __shfl(wave_sum_f64(s + c), 63, 64);
      if (lane == 0) red[wave] = tot;
      __syncthreads();
      if (threadIdx.x == 0) {
        double e = 0.0, ec = 0.0;
        for (uint32_t w = 0; w < W; ++w) dd_acc(e, ec, red[w]);
        J.acc[0] = e;
        J.acc[1] = ec;
      }
      __syncthreads();
    }
    for (uint32_t p = 0; p < J.nrev; ++p) {
      const QC PassArgs& a = *(const QC PassArgs*)(J.rev + p);
      for (uint64_t wt = (uint64_t)wave; wt < (dim >> 9); wt += W)
        adjoint_tile(a, wt, psi, lam, xbuf[wave], lane, [&](uint32_t g, double tot) {
          if (lane == 0) pts[g][wave] += tot;
        });
      __syncthreads();
      if (threadIdx.x == 0) {
        const uint32_t off = J.map_off[p], cnt = J.map_cnt[p];
        for (uint32_t g = 0; g < cnt; ++g) {
          double v = 0.0;
          for (uint32_t w = 0; w < W; ++w) {
            v += pts[g][w];
            pts[g][w] = 0.0;
          }
          double* acc = J.acc + 2 + 2 * (uint64_t)J.pmap[off + g];
          double hi = acc[0], lo = acc[1];
          dd_acc(hi, lo, v);
          acc[0] = hi;
          acc[1] = lo;
        }
      }
      __syncthreads();
    }
  }
}

hipError_t launch_batch_persistent(const BatchJob* jobs, uint32_t njobs, unsigned* next, unsigned groups,
                                   hipStream_t st) {
  if (njobs == 0 || groups == 0) return hipSuccess;
  hipLaunchKernelGGL(batch_persistent_kernel, dim3(groups), dim3(kThreads), 0, st, jobs, njobs, next);
  return hipGetLastError();
}

// Paired form on the forward engine's 1024-amplitude tiles: a workgroup of
// two waves holds one tile of psi (wave 0) and the same positions of lambda
// (wave 1), 16 rows each, and both waves run every dense / elementwise op.
// At the first of a run of adjacent gradient points the psi wave parks its
// tile in LDS, and for every point of the run the lambda wave reads the Pauli
// partner of each element from there, whatever the flip mask.  Two LDS tiles
// alternate between runs, so one barrier per run keeps the psi wave from
// overwriting a tile the lambda wave still reads.
template <int KV>
__global__ __launch_bounds__(128, KV == 16 ? 2 : 3) void wave_pair_adjoint_kernel(const PassArgs* __restrict__ pa,
                                                                   double2* __restrict__ psi,
                                                                   double2* __restrict__ lam,
                                                                   double* __restrict__ partials, uint64_t tile0,
                                                                   uint64_t ntiles) {
  constexpr int RB = KV == 16 ? 4 : 3;
  __shared__ __align__(16) double2 pbuf[2][64 * KV];
  // the lambda wave's Walsh transform of a run's Z-only products, row z at 64 z
  __shared__ __align__(16) double wh[64 * KV];
  const QC PassArgs& a = *(const QC PassArgs*)(pa);
  const QC Op* ops = (const QC Op*)(a.ops);
  const QC PhaseTerm* terms = (const QC PhaseTerm*)(a.terms);
  const QC DiagTable* diags = (const QC DiagTable*)(a.diags);
  const QC double* coefs = (const QC double*)(a.coefs);
  const int lane = threadIdx.x & 63;
  const bool is_lam = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) != 0;
  const uint64_t wt = tile0 + blockIdx.x;
  if (wt >= ntiles) return;  // uniform over the workgroup
  const TileMapConst tm{(uint32_t)(6 + RB), 6u, (uint32_t)RB, a.E};
  const uint64_t base = tm.base(wt);
  const uint64_t e0 = 1ull << a.E[0], e1 = 1ull << a.E[1], e2 = 1ull << a.E[2];
  const uint64_t e3 = RB > 3 ? 1ull << a.E[3] : 0ull;
  auto roff = [&](int v) -> uint64_t {
    return ((v & 1) ? e0 : 0ull) | ((v & 2) ? e1 : 0ull) | ((v & 4) ? e2 : 0ull) | ((v & 8) ? e3 : 0ull);
  };
  const WaveCtx cx{lane, base, wt, a.ptab};
  QG double* sv = (QG double*)(is_lam ? lam : psi) + 2 * (base + lane);
  double2 amp[KV];
#pragma unroll
  for (int v = 0; v < KV; ++v) amp[v] = make_double2(sv[2 * roff(v)], sv[2 * roff(v) + 1]);
  uint32_t pb = 0;
  for (uint32_t o = a.op_begin; o < a.op_end;) {
    const QC Op& op = ops[o];
    if (op.kind == OP_DENSE) {
      wave_dense(op, amp, cx, coefs);
      ++o;
      continue;
    }
    if (op.kind == OP_ELEM) {
      wave_elem(op, amp, cx, terms, diags, coefs);
      ++o;
      continue;
    }
    // a run of adjacent points sees one psi tile, parked once.  The whole run
    // is taken in this iteration: the points leave the amplitudes as they
    // are, so a run costs one pass through the loop's latch (where the
    // amplitudes return to the registers the loop header holds them in)
    // instead of one per point.
    double2* xb = pbuf[pb];
    pb ^= 1u;
    if (!is_lam) {
#pragma unroll
      for (int v = 0; v < KV; ++v) xb[64 * v + lane] = amp[v];
    }
    __syncthreads();
    if (!is_lam) {
      while (o < a.op_end && ops[o].kind == OP_GRAD) ++o;
      continue;
    }
    // point P: Re(i^kr conj(l) x) over the tile, x the psi partner at index
    // q = (64 v + lane) ^ fx, times the Z/Y sign (-1)^popc(q & zt).  With r
    // and lf the row and lane parts of fx and zr, zl those of zt, that sign
    // is (-1)^(popc(v & zr) + popc(r & zr) + popc((lane ^ lf) & zl)): the
    // first factor is a Walsh fold of the row products, the second uniform,
    // the third per lane.  i^kr picks Re(conj(l) x) or Im(conj(l) x) and an
    // overall sign.
    bool hat = false;  // wh holds this run's transform
    for (; o < a.op_end && ops[o].kind == OP_GRAD; ++o) {
      const QC Op& pt = ops[o];
      const uint32_t fx = (uint32_t)pt.gfx, zt = (uint32_t)pt.gzt;
      const uint32_t kr = (1u + pt.nY) & 3u;  // i * i^nY = i^kr
      const uint32_t r = (fx >> 6) & (uint32_t)(KV - 1), lf = fx & 63u;
      const uint32_t zr = (zt >> 6) & (uint32_t)(KV - 1), zl = zt & 63u;
      double w;
      if (fx == 0u && (kr & 1u)) {
        // a Z-only point (no flips, so nY = 0 and kr = 1): its row fold is
        // entry zr of the Walsh transform of the run's Im(conj(l) x)
        // products, formed at the run's first such point; each lane reads
        // back only what it wrote
        if (!hat) {
          double t[KV];
          pair_point_products<KV, 0>(xb + lane, amp, true, t);
          walsh_transform<KV>(t);
#pragma unroll
          for (int v = 0; v < KV; ++v) wh[64 * v + lane] = t[v];
          hat = true;
        }
        w = wh[64 * zr + (uint32_t)lane];
      } else {
        double t[KV];
        pair_point_products_r<KV>(xb + ((uint32_t)lane ^ lf), amp, r, (kr & 1u) != 0u, t);
        w = walsh_fold<KV>(t, zr);
      }
      const bool lneg = __popc(((uint32_t)lane ^ lf) & zl) & 1;
      const bool flip_all =
          (__popcll(base & pt.gzb) & 1) ^ (kr == 1u || kr == 2u) ^ (__popc(r & zr) & 1);
      // tile sum: the lanes' sums added in a DPP tree
      double tot = __shfl(wave_sum_f64(lneg ? -w : w), 63, 64);
      tot = flip_all ? -tot : tot;
      if (lane == 0) {
        partials[((uint64_t)pt.gidx * ntiles + wt) * 2] = tot;
        partials[((uint64_t)pt.gidx * ntiles + wt) * 2 + 1] = 0.0;
      }
    }
  }
#pragma unroll
  for (int v = 0; v < KV; ++v) {
    sv[2 * roff(v)] = amp[v].x;
    sv[2 * roff(v) + 1] = amp[v].y;
  }
}

// Phase tables of one elementwise op (layout beside kPtabC in
// qsim_internal.h): one entry per thread, the angles of the op's matching
// terms summed in term order and turned into cis().
__global__ __launch_bounds__(kThreads) void phase_table_kernel(const PassArgs* __restrict__ pa,
                                                               const Op* __restrict__ opp, [[maybe_unused]] uint64_t ntau,
                                                               uint64_t total, double2* __restrict__ ptab) {
  const PassArgs& a = *pa;
  const Op& op = *opp;
  const uint64_t e = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
  if (e >= total) return;
  const TileMapDev tm{a.t, 6u, a.t - 6u, a.E};
  enum { K_C, K_R, K_F, K_Q, K_U, K_G, K_M } kind;
  uint32_t lane = 0, j = 0, v = 0;
  uint64_t base = 0;
  const uint64_t qo = op.ptab_q, uo = qo + 64, go = op.ptab_g, mo = op.ptab_m;
  if (e < kPtabR) {
    kind = K_C;
    j = (uint32_t)(e >> 6);
    lane = (uint32_t)(e & 63u);
  } else if (e < kPtabF) {
    kind = K_R;
    v = (uint32_t)(e - kPtabR);
  } else if (e < qo) {
    kind = K_F;
    j = (uint32_t)((e - kPtabF) & 3u);
    base = tm.base((e - kPtabF) >> 2);
  } else if (e < uo) {
    kind = K_Q;
    lane = (uint32_t)(e - qo);
  } else if (e < go) {
    kind = K_U;
    base = tm.base(e - uo);
  } else if (e < mo) {
    kind = K_G;
    j = (uint32_t)((e - go) & 7u);  // the lane bit
    base = tm.base((e - go) >> 3);
  } else {
    kind = K_M;
    lane = (uint32_t)((e - mo) & 63u);
    const uint64_t row = (e - mo) >> 6;
    if (op.ptab_flags & kPtabK) {
      // the row is the value of the key bits; the terms of K test no others
      for (uint32_t i = 0; i < op.ptab_nk; ++i) base |= ((row >> i) & 1ull) << op.ptab_kpos[i];
    } else {
      base = tm.base(row);
    }
  }
  double s = 0.0;
  for (uint32_t i = 0; i < op.nterms; ++i) {
    const PhaseTerm& T = a.terms[op.term_off + i];
    if (T.bval & ~T.bmask) continue;  // padding: never matches
    const uint32_t lm = T.tmask & 63u, lv = T.tval & 63u;
    const uint32_t rm = (T.tmask >> 6) & 15u, rv = (T.tval >> 6) & 15u;
    const bool lane_ok = (lane & lm) == lv, base_ok = (base & T.bmask) == T.bval;
    // the slot-free part of a term (rm == 0, or a single-slot term's s_j = 0
    // share) goes to Q, U or M by its conditions
    const bool slot_free = rm == 0u || (!(rm & (rm - 1u)) && rv == 0u);
    if (slot_free) {
      if ((op.ptab_flags & kPtabG) && lm && T.bmask && !(lm & (lm - 1u))) {
        // one lane bit k and a base condition: G[tau][k] (+angle for bit
        // k = 1; for bit k = 0 +angle to U and -angle to G)
        if (kind == K_G && (uint32_t)(__ffs(lm) - 1) == j && base_ok) s += lv ? T.angle : -T.angle;
        if (kind == K_U && !lv && base_ok) s += T.angle;
      } else {
        const int fk = (op.ptab_flags & kPtabM) ? K_M : (lm && T.bmask) ? K_M : (T.bmask ? K_U : K_Q);
        if (kind == fk && lane_ok && base_ok) s += T.angle;
      }
    }
    if (rm & (rm - 1u)) {
      if (kind == K_R && (v & rm) == rv) s += T.angle;
    } else if (rm) {
      if ((uint32_t)(__ffs(rm) - 1) == j) {
        const double sa = rv ? T.angle : -T.angle;
        if (kind == K_C && T.bmask == 0ull && lane_ok) s += sa;
        if (kind == K_F && T.bmask != 0ull && base_ok) s += sa;
      }
    }
  }
  // G entries stay angles (the kernel adds a lane's angles, then takes cis)
  ptab[op.ptab_off + e] = kind == K_G ? make_double2(s, 0.0) : cis_reduced(s);
}

hipError_t launch_phase_table(const PassArgs* d_args, const Op* d_op, uint64_t ntau, uint64_t entries, double* ptab,
                              hipStream_t st) {
  const uint64_t groups = (entries + kThreads - 1) / kThreads;
  if (groups == 0) return hipSuccess;
  if (groups > kMaxLaunchBlocks) return hipErrorInvalidValue;
  hipLaunchKernelGGL(phase_table_kernel, dim3((unsigned)groups), dim3(kThreads), 0, st, d_args, d_op, ntau, entries,
                     reinterpret_cast<double2*>(ptab));
  return hipGetLastError();
}

// ntiles counts tiles of 2^tile_bits amplitudes (10: 16 rows per wave, 9: 8
// rows) of one state; one workgroup per tile
hipError_t launch_wave_pair_adjoint(int tile_bits, const PassArgs* d_args, double2* psi, double2* lam,
                                    double* partials, uint64_t ntiles, hipStream_t st) {
  if (tile_bits != 10 && tile_bits != 9) return hipErrorInvalidValue;
  const uint64_t maxg = max_launch_groups();
  for (uint64_t t0 = 0; t0 < ntiles; t0 += maxg) {
    const uint64_t cnt = std::min<uint64_t>(ntiles - t0, maxg);
    if (tile_bits == 10)
      hipLaunchKernelGGL(wave_pair_adjoint_kernel<16>, dim3((unsigned)cnt), dim3(128), 0, st, d_args, psi, lam,
                         partials, t0, ntiles);
    else
      hipLaunchKernelGGL(wave_pair_adjoint_kernel<8>, dim3((unsigned)cnt), dim3(128), 0, st, d_args, psi, lam,
                         partials, t0, ntiles);
  }
  return hipGetLastError();
}

// ntiles counts wave tiles of 2^9 amplitudes of one state
hipError_t launch_wave_adjoint(int tile_bits, const PassArgs* d_args, double2* psi, double2* lam, double* partials,
                               uint64_t ntiles, hipStream_t st) {
  constexpr uint64_t per_block = kThreads / 64;
  if (tile_bits != 9) return hipErrorInvalidValue;
  const uint64_t maxg = max_launch_groups();
  for (uint64_t t0 = 0; t0 < ntiles; t0 += maxg * per_block) {
    const uint64_t cnt = std::min<uint64_t>(ntiles - t0, maxg * per_block);
    hipLaunchKernelGGL(wave_adjoint_kernel, dim3((unsigned)((cnt + per_block - 1) / per_block)), dim3(kThreads), 0,
                       st, d_args, psi, lam, partials, t0, ntiles);
  }
  return hipGetLastError();
}

uint64_t wave_persistent_waves(uint64_t ntiles) { return persistent_waves(ntiles); }

// Forms that decode, and RAW->RAW on 1024-amplitude tiles: one launch, a
// persistent grid of at most persistent_groups() work-groups (each wave walks
// its tiles with a stride of the grid).  RAW->ENC, and RAW->RAW on
// 512-amplitude tiles: one tile per wave, launches split at kMaxLaunchBlocks
// work-groups.
template <int IN, int OUT, int KV, bool PT = false>
static hipError_t launch_wave_mode(const PassArgs* d_args, const PassDyn& dy, uint64_t ntiles, hipStream_t st) {
  constexpr uint64_t per_block = kThreads / 64;
  if constexpr (IN == IN_DEC || (IN == IN_RAW && OUT == OUT_RAW && KV == 16)) {
    static_assert(per_block == 4, "persistent_waves assumes four waves per work-group");
    const uint64_t groups = persistent_waves(ntiles) / per_block;
    if (groups == 0) return hipSuccess;
    hipLaunchKernelGGL((wave_kernel<IN, OUT, KV, PT>), dim3((unsigned)groups), dim3(kThreads), 0, st, d_args, dy,
                       dy.tile0 + ntiles);
  } else {
    const uint64_t maxb = max_launch_groups();
    PassDyn d = dy;
    for (uint64_t t0 = 0; t0 < ntiles; t0 += maxb * per_block) {
      const uint64_t cnt = std::min<uint64_t>(ntiles - t0, maxb * per_block);
      d.tile0 = dy.tile0 + t0;
      hipLaunchKernelGGL((wave_kernel<IN, OUT, KV, PT>), dim3((unsigned)((cnt + per_block - 1) / per_block)),
                         dim3(kThreads), 0, st, d_args, d, d.tile0 + cnt);
    }
  }
  return hipGetLastError();
}

template <int KV>
static hipError_t launch_wave_kv(int in_mode, int out_mode, const PassArgs* d_args, const PassDyn& dy, uint64_t ntiles,
                                 hipStream_t st) {
  if (in_mode == IN_RAW && out_mode == OUT_RAW) return launch_wave_mode<IN_RAW, OUT_RAW, KV>(d_args, dy, ntiles, st);
  if (in_mode == IN_DEC && out_mode == OUT_RAW) return launch_wave_mode<IN_DEC, OUT_RAW, KV>(d_args, dy, ntiles, st);
  if (in_mode == IN_RAW && out_mode == OUT_ENC) return launch_wave_mode<IN_RAW, OUT_ENC, KV>(d_args, dy, ntiles, st);
  return launch_wave_mode<IN_DEC, OUT_ENC, KV>(d_args, dy, ntiles, st);
}

// A lambda-only reverse pass on 1024-amplitude tiles (dy.raw: lambda,
// dy.psi: the stored forward state, dy.partials: the points' tile sums).
hipError_t launch_wave_lam(const PassArgs* d_args, const PassDyn& dy, uint64_t ntiles, hipStream_t st) {
  return launch_wave_mode<IN_RAW, OUT_RAW, 16, true>(d_args, dy, ntiles, st);
}

// ntiles counts wave tiles (2^tile_bits amplitudes each, tile_bits 10 or 9)
// of this launch
hipError_t launch_wave(int tile_bits, int in_mode, int out_mode, const PassArgs* d_args, const PassDyn& dy,
                       uint64_t ntiles, hipStream_t st) {
  if (tile_bits == 10) return launch_wave_kv<16>(in_mode, out_mode, d_args, dy, ntiles, st);
  if (tile_bits == 9) return launch_wave_kv<8>(in_mode, out_mode, d_args, dy, ntiles, st);
  return hipErrorInvalidValue;
}

}  // namespace qsim
