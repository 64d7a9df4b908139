// Wave-resident pass kernel for gfx950 (CDNA4, wave64).
//
// Each wave owns a "wave tile" of 1024 amplitudes of the group-local state:
// tile bits 0..5 are the lane (one complex128 per lane, a 1 KiB coalesced row;
// each half-wave holds one 32-amplitude codec chunk) and tile bits 6..9 are
// four register slots (16 complex128 per lane in VGPRs) placed on the pass's
// register qubits E[0..3].  Gates on lane bits exchange partners across lanes,
// gates on register slots are in-lane, and runs of diagonal gates become
// per-lane and per-slot phase factors.  There is no LDS tile and no
// work-group barrier: waves progress independently, so occupancy is bounded by
// VGPRs only.  The codec works per wave tile: half-wave DPP reductions for the
// chunk exponent and width, one allocation atomic per wave tile, and wave-local
// LDS staging for the bit packing.  Results are bit-identical to the LDS-tile
// kernel (same chunk format and quantiser).
#include "qsim_device.h"

namespace qsim {

constexpr int kRegBits = 4;
constexpr int kV = 1 << kRegBits;  // complex values per lane
constexpr int kWaveTileBits = 6 + kRegBits;

__device__ __forceinline__ uint32_t half_bits(uint64_t bal, int lane) {
  return lane < 32 ? (uint32_t)bal : (uint32_t)(bal >> 32);
}

// 32-lane reductions: lanes 0..31 get the reduction of lanes 0..31, lanes
// 32..63 that of lanes 32..63.  All 64 lanes must be active.
__device__ __forceinline__ int half_min_i32(int v, int lane) {
  const int id = 0x7fffffff;
  v = min(v, __builtin_amdgcn_update_dpp(id, v, 0xb1, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(id, v, 0x4e, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(id, v, 0x141, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(id, v, 0x140, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(id, v, 0x142, 0xa, 0xf, false));
  const int a = __builtin_amdgcn_readlane(v, 31), b = __builtin_amdgcn_readlane(v, 63);
  return lane < 32 ? a : b;
}
__device__ __forceinline__ int half_max_i32(int v, int lane) {
  const int id = (int)0x80000000;
  v = max(v, __builtin_amdgcn_update_dpp(id, v, 0xb1, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(id, v, 0x4e, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(id, v, 0x141, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(id, v, 0x140, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(id, v, 0x142, 0xa, 0xf, false));
  const int a = __builtin_amdgcn_readlane(v, 31), b = __builtin_amdgcn_readlane(v, 63);
  return lane < 32 ? a : b;
}

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
// keeps the compiler from hoisting coefficient loads of later register
// iterations (bounds VGPR live ranges)
#define QSIM_ITER_FENCE() asm volatile("" ::: "memory")

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

// ------------------------------------------------------------------ codec
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

// Doubles 2i and 2i+1 of a chunk (the lane's re and im).  For w <= 32 the
// 2w-bit field of lane i starts at bit 2iw; three dwords from dword
// floor(2iw/32) always cover it (the arena keeps 8 bytes of padding past its
// last payload, so the third dword is always mapped).
__device__ __forceinline__ double2 decode_pair(uint64_t d, const QG uint64_t* __restrict__ arena, uint32_t i) {
  const uint32_t w = desc_w(d);
  const int e = desc_e(d);
  const QG uint64_t* P = arena + desc_off(d);
  double2 r = make_double2(0.0, 0.0);
  if (w - 1u < 32u) {
    const QG uint32_t* P32 = (const QG uint32_t*)(P);
    const uint32_t bit = 2u * i * w, q = bit >> 5, s = bit & 31u;
    const uint32_t d0 = P32[q], d1 = P32[q + 1], d2 = P32[q + 2];
    const uint32_t lo = __builtin_amdgcn_alignbit(d1, d0, s), hi = __builtin_amdgcn_alignbit(d2, d1, s);
    const uint64_t x = ((uint64_t)hi << 32) | lo;
    const uint64_t m = (2ull << (w - 1u)) - 1ull;
    const uint32_t u0 = (uint32_t)(x & m), u1 = (uint32_t)((x >> w) & m);
    const int k0 = (int)(u0 >> 1) ^ -(int)(u0 & 1u), k1 = (int)(u1 >> 1) ^ -(int)(u1 & 1u);
    r = make_double2(ldexp((double)k0, e), ldexp((double)k1, e));
  } else if (w == 64u) {
    r = make_double2(__longlong_as_double((long long)P[2 * i]), __longlong_as_double((long long)P[2 * i + 1]));
  } else if (w != 0u) {
    r = make_double2(unzig_ldexp(extract64(P, 2u * i * w, w), e), unzig_ldexp(extract64(P, (2u * i + 1u) * w, w), e));
  }
  return r;
}

// OR over each half-wave: lane 31 ends with lanes 0..31, lane 63 with 32..63
__device__ __forceinline__ uint32_t half_or_u32(uint32_t v) {
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xb1, 0xf, 0xf, false);
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4e, 0xf, 0xf, false);
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xf, 0xf, false);
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xf, 0xf, false);
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);
  return v;
}

struct HalfPlan {
  uint32_t a, b;  // w | (e & 0xffff) << 8 of the low and the high half-wave
};

// Quantise the two chunks held by the wave (half-wave h holds the doubles
// x0, x1 of its 32 lanes) and replace x0, x1 by their codes: zig-zag integers
// of rint(x * 2^-e), or the raw bit patterns when a chunk is stored verbatim
// (w = 64).  e is the minimum over nonzero doubles of
// ilogb(2*fma(eps_rel,|x|,eps_abs)) minus the guard bits; w the largest code
// bit length; an all-zero chunk gets w = e = 0.  rint uses the 1.5*2^52 shift
// (exact round-to-nearest-even for |s| < 2^51, the bound that selects the
// quantised form).  Every lane of the wave must call it.
__device__ __forceinline__ HalfPlan half_quantise(double& x0, double& x1, const CodecParams& cp, int lane) {
  const bool nz = x0 != 0.0 || x1 != 0.0;
  const uint64_t bnz = ballot64(nz);
  const int le = min(lane_exponent(x0, cp), lane_exponent(x1, cp));
  const int r = half_min_i32(le, lane);
  const int ra = __builtin_amdgcn_readlane(r, 0), rb = __builtin_amdgcn_readlane(r, 32);
  const bool rawA = ra == (int)0x80000000 || ra < -1000, rawB = rb == (int)0x80000000 || rb < -1000;
  const int eA = rawA ? 0 : ra - cp.guard, eB = rawB ? 0 : rb - cp.guard;
  const int e = lane < 32 ? eA : eB;
  const double s0 = ldexp(x0, -e), s1 = ldexp(x1, -e);
  const uint64_t bbad = ballot64(!(fabs(s0) < 2251799813685248.0) || !(fabs(s1) < 2251799813685248.0));
  const bool vA = !rawA && (uint32_t)bbad == 0u, vB = !rawB && (uint32_t)(bbad >> 32) == 0u;
  const double M = 6755399441055744.0;  // 1.5 * 2^52
  const int64_t k0 = __double_as_longlong(s0 + M) - 0x4338000000000000ll;
  const int64_t k1 = __double_as_longlong(s1 + M) - 0x4338000000000000ll;
  const uint64_t u0 = ((uint64_t)k0 << 1) ^ (uint64_t)(k0 >> 63);
  const uint64_t u1 = ((uint64_t)k1 << 1) ^ (uint64_t)(k1 >> 63);
  const uint64_t m = u0 | u1;
  const uint32_t olo = half_or_u32((uint32_t)m), ohi = half_or_u32((uint32_t)(m >> 32));
  const uint64_t orA = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)ohi, 31) << 32) |
                       (uint32_t)__builtin_amdgcn_readlane((int)olo, 31);
  const uint64_t orB = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)ohi, 63) << 32) |
                       (uint32_t)__builtin_amdgcn_readlane((int)olo, 63);
  const uint32_t wA = orA ? 64u - (uint32_t)__clzll(orA) : 0u, wB = orB ? 64u - (uint32_t)__clzll(orB) : 0u;
  HalfPlan hp;
  hp.a = (uint32_t)bnz == 0u ? 0u : (!vA ? 64u : (wA | ((uint32_t)(eA & 0xffff) << 8)));
  hp.b = (uint32_t)(bnz >> 32) == 0u ? 0u : (!vB ? 64u : (wB | ((uint32_t)(eB & 0xffff) << 8)));
  const bool verbatim = lane < 32 ? !vA : !vB;
  x0 = verbatim ? x0 : __longlong_as_double((long long)u0);
  x1 = verbatim ? x1 : __longlong_as_double((long long)u1);
  return hp;
}

// OR a code of nbits (<= 64) at bit offset `bit` into the wave's staging row
__device__ __forceinline__ void stage_or(uint64_t* stg, uint32_t bit, uint32_t nbits, uint64_t c) {
  if (c == 0) return;
  const uint32_t q = bit >> 6, s = bit & 63u;
  atomicOr(reinterpret_cast<unsigned long long*>(&stg[q]), (unsigned long long)(c << s));
  if (s + nbits > 64u) atomicOr(reinterpret_cast<unsigned long long*>(&stg[q + 1]), (unsigned long long)(c >> (64u - s)));
}

// ------------------------------------------------------------ dense ops
struct WaveCtx {
  int lane;
  uint64_t base;
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

// one target on lane bit b
__device__ __forceinline__ void k1_lane(const QC Op& op, double2 (&amp)[kV], const WaveCtx& cx, const QC double* __restrict__ C) {
  const uint32_t b = op.tpos[0];
  const uint32_t r = ((uint32_t)cx.lane >> b) & 1u;
  const int m = 1 << b;
  if (op.nsel == 0) {
    if (op.ident_mask & 1ull) return;
    const double2 md = ldc(C, op.mat_off + 3u * r), mo = ldc(C, op.mat_off + 1u + r);
#pragma unroll
    for (int v = 0; v < kV; ++v) {
      const double2 p = shfl_xor2(amp[v], m);
      amp[v] = mac2(md, amp[v], mo, p);
      QSIM_ITER_FENCE();
    }
    return;
  }
#pragma unroll
  for (int v = 0; v < kV; ++v) {
    QSIM_ITER_FENCE();
    const double2 p = shfl_xor2(amp[v], m);
    const uint32_t sel = wave_sel(op, cx, v);
    if ((op.ident_mask >> sel) & 1ull) continue;
    const uint64_t mo0 = op.mat_off + 4ull * sel;
    amp[v] = mac2(ldc(C, mo0 + 3u * r), amp[v], ldc(C, mo0 + 1u + r), p);
  }
}

// one target on register slot J
template <int J>
__device__ __forceinline__ void k1_reg(const QC Op& op, double2 (&amp)[kV], const WaveCtx& cx, const QC double* __restrict__ C) {
  constexpr int MJ = 1 << J;
#pragma unroll
  for (int v = 0; v < kV; ++v) {
    if (v & MJ) continue;
    QSIM_ITER_FENCE();
    const uint32_t sel = op.nsel ? wave_sel(op, cx, v) : 0u;
    if ((op.ident_mask >> sel) & 1ull) continue;
    const uint64_t mo = op.mat_off + 4ull * sel;
    const double2 a = amp[v], b = amp[v | MJ];
    amp[v] = mac2(ldc(C, mo), a, ldc(C, mo + 1), b);
    amp[v | MJ] = mac2(ldc(C, mo + 2), a, ldc(C, mo + 3), b);
  }
}

// two targets on register slots J0 (matrix bit 0) and J1 (matrix bit 1)
template <int J0, int J1>
__device__ __forceinline__ void k2_regreg(const QC Op& op, double2 (&amp)[kV], const WaveCtx& cx, const QC double* __restrict__ C) {
  constexpr int M0 = 1 << J0, M1 = 1 << J1;
#pragma unroll
  for (int v = 0; v < kV; ++v) {
    if (v & (M0 | M1)) continue;
    QSIM_ITER_FENCE();
    const uint32_t sel = op.nsel ? wave_sel(op, cx, v) : 0u;
    if ((op.ident_mask >> sel) & 1ull) continue;
    const uint64_t mo = op.mat_off + 16ull * sel;
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

// two targets: lane bit b (matrix bit il) and register slot J (matrix bit 1-il)
template <int J>
__device__ __forceinline__ void k2_lanereg(const QC Op& op, double2 (&amp)[kV], const WaveCtx& cx, const QC double* __restrict__ C,
                                           uint32_t b, uint32_t il) {
  constexpr int MJ = 1 << J;
  const uint32_t ir = 1u - il;
  const uint32_t rl = ((uint32_t)cx.lane >> b) & 1u;
  const int m = 1 << b;
#pragma unroll
  for (int v = 0; v < kV; ++v) {
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
__device__ __forceinline__ void k2_lanelane(const QC Op& op, double2 (&amp)[kV], const WaveCtx& cx, const QC double* __restrict__ C) {
  const uint32_t b0 = op.tpos[0], b1 = op.tpos[1];
  const uint32_t r = (((uint32_t)cx.lane >> b0) & 1u) | ((((uint32_t)cx.lane >> b1) & 1u) << 1);
#pragma unroll
  for (int v = 0; v < kV; ++v) {
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

__device__ __forceinline__ void wave_dense(const QC Op& op, double2 (&amp)[kV], const WaveCtx& cx, const QC double* __restrict__ C) {
  if (op.k == 1) {
    const uint32_t t = op.tpos[0];
    if (t < 6) k1_lane(op, amp, cx, C);
    else switch (t - 6) {
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
__device__ __forceinline__ void wave_elem(const QC Op& op, double2 (&amp)[kV], const WaveCtx& cx,
                                          const QC PhaseTerm* __restrict__ terms, const QC DiagTable* __restrict__ diags,
                                          const QC double* __restrict__ C) {
  const uint32_t lane = (uint32_t)cx.lane;
  if (op.nterms) {
    double aL = 0.0, aF0 = 0.0, aF1 = 0.0, aF2 = 0.0, aF3 = 0.0;
    uint32_t used = 0;
    bool generic = false;
    for (uint32_t i = 0; i < op.nterms; ++i) {
      const QC PhaseTerm& T = terms[op.term_off + i];
      if ((cx.base & T.bmask) != T.bval) continue;
      const uint32_t lm = T.tmask & 63u, lv = T.tval & 63u;
      const uint32_t rm = (T.tmask >> 6) & 15u, rv = (T.tval >> 6) & 15u;
      const double a = ((lane & lm) == lv) ? T.angle : 0.0;
      if (rm == 0) { aL += a; used |= 16u; continue; }
      if (rm & (rm - 1u)) { generic = true; continue; }
      const uint32_t j = __ffs(rm) - 1;
      double f = a;
      if (!rv) { aL += a; f = -a; used |= 16u; }
      aF0 += j == 0 ? f : 0.0;
      aF1 += j == 1 ? f : 0.0;
      aF2 += j == 2 ? f : 0.0;
      aF3 += j == 3 ? f : 0.0;
      used |= 1u << j;
    }
    double2 L = make_double2(1.0, 0.0), F0 = L, F1 = L, F2 = L, F3 = L;
    if (used & 16u) sincos(aL, &L.y, &L.x);
    if (used & 1u) sincos(aF0, &F0.y, &F0.x);
    if (used & 2u) sincos(aF1, &F1.y, &F1.x);
    if (used & 4u) sincos(aF2, &F2.y, &F2.x);
    if (used & 8u) sincos(aF3, &F3.y, &F3.x);
    if (used & 16u) {
#pragma unroll
      for (int v = 0; v < kV; ++v) amp[v] = cmul(amp[v], L);
    }
#pragma unroll
    for (int v = 0; v < kV; ++v) {
      if ((v & 1) && (used & 1u)) amp[v] = cmul(amp[v], F0);
      if ((v & 2) && (used & 2u)) amp[v] = cmul(amp[v], F1);
      if ((v & 4) && (used & 4u)) amp[v] = cmul(amp[v], F2);
      if ((v & 8) && (used & 8u)) amp[v] = cmul(amp[v], F3);
    }
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
        double sn, cs;
        sincos(T.angle, &sn, &cs);
        const double2 f = ((lane & lm) == lv) ? make_double2(cs, sn) : make_double2(1.0, 0.0);
#pragma unroll
        for (int v = 0; v < kV; ++v)
          if (((uint32_t)v & rm) == rv) amp[v] = cmul(amp[v], f);
      }
    }
  }
  for (uint32_t d = 0; d < op.ndiag; ++d) {
    const QC DiagTable& Dt = diags[op.diag_off + d];
#pragma unroll
    for (int v = 0; v < kV; ++v) {
      uint32_t idx = 0;
      for (uint32_t j = 0; j < Dt.m; ++j) {
        const uint32_t pos = Dt.pos[j];
        uint32_t bit;
        if (!Dt.tile[j]) bit = (uint32_t)((cx.base >> pos) & 1ull);
        else if (pos < 6) bit = (lane >> pos) & 1u;
        else bit = ((uint32_t)v >> (pos - 6)) & 1u;
        idx |= bit << j;
      }
      amp[v] = cmul(amp[v], ldc(C, Dt.coef_off + idx));
    }
  }
}

// ------------------------------------------------------------- the kernel
template <int IN, int OUT>
__global__ __launch_bounds__(kThreads, 2) void wave_kernel(const PassArgs* __restrict__ pa, PassDyn dy, uint64_t ntiles) {
  // per-wave staging for the packed payload of half a wave tile (8 register
  // rows = 16 chunks of at most 64 words), plus 8 words of slack for the
  // zero-valued spill OR past the last field
  __shared__ uint64_t stg[kThreads / 64][kV / 2 * 2 * 64 + 8];
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
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t wt = dy.tile0 + (uint64_t)blockIdx.x * (kThreads / 64) + wave;
  if (wt >= ntiles) return;
  const uint32_t L = a.L;
  const uint64_t slot = wt >> (L - kWaveTileBits);
  const uint64_t tau = wt & ((1ull << (L - kWaveTileBits)) - 1ull);
  const TileMapConst tm{kWaveTileBits, 6u, (uint32_t)kRegBits, a.E};
  const uint64_t base = tm.base(tau);
  const uint64_t e0 = 1ull << a.E[0], e1 = 1ull << a.E[1], e2 = 1ull << a.E[2], e3 = 1ull << a.E[3];
  auto roff = [&](int v) -> uint64_t {
    return ((v & 1) ? e0 : 0ull) | ((v & 2) ? e1 : 0ull) | ((v & 4) ? e2 : 0ull) | ((v & 8) ? e3 : 0ull);
  };
  const WaveCtx cx{lane, base};
  double2 amp[kV];

  // chunk metadata: lane i < 2*kV describes chunk i = (register index i>>1, half i&1)
  uint64_t mdesc = 0;
  uint32_t mbid = 0, mci = 0, mnull = 1;
  if constexpr (IN == IN_DEC || OUT == OUT_ENC) {
    if (lane < 2 * kV) {
      const uint32_t p0 = ((uint32_t)(lane >> 1) << 6) | ((uint32_t)(lane & 1) << 5);
      const uint64_t loc = base | tm.off(p0);
      mbid = bmap[(slot << a.khi) + (loc >> a.B)];
      mci = (uint32_t)((loc & ((1ull << a.B) - 1ull)) >> kChunkAmpsLog2);
      mnull = null_in[mbid];
      if constexpr (IN == IN_DEC) mdesc = mnull ? 0ull : desc_in[(uint64_t)mbid * a.cpb + mci];
    }
  }

  // ---- load
  if constexpr (IN == IN_RAW) {
    const QG double* src = (const QG double*)(dy.raw) + 2 * ((slot << L) + base + lane);
#pragma unroll
    for (int v = 0; v < kV; ++v) amp[v] = make_double2(src[2 * roff(v)], src[2 * roff(v) + 1]);
  } else {
#pragma unroll
    for (int v = 0; v < kV; ++v) {
      const uint64_t dA = readlane64(mdesc, 2 * v), dB = readlane64(mdesc, 2 * v + 1);
      amp[v] = decode_pair(lane < 32 ? dA : dB, arena_in, (uint32_t)lane & 31u);
    }
  }

  // ---- ops
  for (uint32_t o = a.op_begin; o < a.op_end; ++o) {
    const QC Op& op = ops[o];
    if (op.kind == OP_DENSE) wave_dense(op, amp, cx, coefs);
    else wave_elem(op, amp, cx, terms, diags, coefs);
  }

  // ---- store
  if constexpr (OUT == OUT_RAW) {
    QG double* dst = (QG double*)(dy.raw) + 2 * ((slot << L) + base + lane);
#pragma unroll
    for (int v = 0; v < kV; ++v) {
      dst[2 * roff(v)] = amp[v].x;
      dst[2 * roff(v) + 1] = amp[v].y;
    }
  } else {
    const CodecParams cp{a.eps_rel, a.eps_abs, a.guard, a.e_abs, a.p_abs};
    // quantise every register row; chunk j = 2v + h is described by lane j
    uint32_t planA[kV], planB[kV], loffA[kV];
    uint32_t tot = 0, tot_half = 0;
    uint32_t dplan = 0, doff = 0;
#pragma unroll
    for (int v = 0; v < kV; ++v) {
      // zero check of blocks that were null at stage start
      const uint32_t nullA = (uint32_t)__builtin_amdgcn_readlane((int)mnull, 2 * v);
      const uint32_t nullB = (uint32_t)__builtin_amdgcn_readlane((int)mnull, 2 * v + 1);
      if (nullA | nullB) {
        const bool nzv = a.zero_mode == ZERO_EXACT ? (amp[v].x != 0.0 || amp[v].y != 0.0)
                                                   : (fabs(amp[v].x) > a.zero_thr || fabs(amp[v].y) > a.zero_thr);
        const uint64_t bz = ballot64(nzv);
        const uint32_t bidA = (uint32_t)__builtin_amdgcn_readlane((int)mbid, 2 * v);
        const uint32_t bidB = (uint32_t)__builtin_amdgcn_readlane((int)mbid, 2 * v + 1);
        if (lane == 0 && nullA && (uint32_t)bz) nonzero[bidA] = 1;
        if (lane == 32 && nullB && (uint32_t)(bz >> 32)) nonzero[bidB] = 1;
      }
      const HalfPlan hp = half_quantise(amp[v].x, amp[v].y, cp, lane);
      planA[v] = hp.a;
      planB[v] = hp.b;
      if (v == kV / 2) tot_half = tot;
      loffA[v] = tot;
      if ((lane >> 1) == v) {
        dplan = (lane & 1) ? hp.b : hp.a;
        doff = (lane & 1) ? tot + (hp.a & 0xffu) : tot;
      }
      tot += (hp.a & 0xffu) + (hp.b & 0xffu);
    }
    // one allocation per wave tile; its latency is hidden behind the packing
    // of the first half of the rows
    const uint32_t sub = (uint32_t)(wt & dy.sub_mask);
    unsigned long long aoff = 0;
    if (lane == 0 && tot > 0) aoff = atomicAdd(&dy.cursors[sub * 16], (unsigned long long)tot);
    uint64_t* st = stg[wave];
    const uint32_t i = (uint32_t)lane & 31u;
    uint64_t tbase = 0;
    bool skip = false;
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const uint32_t g0 = g == 0 ? 0u : tot_half, g1 = g == 0 ? tot_half : tot;
      for (uint32_t q = lane; q < g1 - g0; q += 64) st[q] = 0ull;
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int v = g * (kV / 2); v < (g + 1) * (kV / 2); ++v) {
        const uint32_t pl = lane < 32 ? planA[v] : planB[v];
        const uint32_t w = pl & 0xffu;
        const uint32_t cb = (lane < 32 ? loffA[v] : loffA[v] + (planA[v] & 0xffu)) - g0;
        const uint64_t u0 = (uint64_t)__double_as_longlong(amp[v].x), u1 = (uint64_t)__double_as_longlong(amp[v].y);
        uint64_t* sc = st + cb;
        if (w - 1u < 32u) {
          const uint64_t c = u0 | (u1 << w);
          const uint32_t bit = 2u * i * w, q = bit >> 6, sh = bit & 63u;
          atomicOr(reinterpret_cast<unsigned long long*>(&sc[q]), (unsigned long long)(c << sh));
          atomicOr(reinterpret_cast<unsigned long long*>(&sc[q + 1]), (unsigned long long)((c >> 1) >> (63u - sh)));
        } else if (w == 64u) {
          sc[2 * i] = u0;
          sc[2 * i + 1] = u1;
        } else if (w != 0u) {
          stage_or(sc, 2u * i * w, w, u0);
          stage_or(sc, (2u * i + 1u) * w, w, u1);
        }
      }
      __builtin_amdgcn_wave_barrier();
      if (g == 0 && tot > 0) {
        const uint64_t off = readlane64(aoff, 0);
        if (off + tot > dy.sub_cap_words) {
          skip = true;
          if (lane == 0) atomicOr(dy.overflow, 1u);
        }
        tbase = off + (uint64_t)sub * dy.sub_cap_words;
      }
      if (!skip)
        for (uint32_t q = lane; q < g1 - g0; q += 64) arena_out[tbase + g0 + q] = st[q];
      __builtin_amdgcn_wave_barrier();
    }
    if (lane < 2 * kV && !skip)
      desc_out[(uint64_t)mbid * a.cpb + mci] = desc_pack(tbase + doff, dplan & 0xffu, (int)(int16_t)(uint16_t)(dplan >> 8));
  }
}

template <int IN, int OUT>
static hipError_t launch_wave_mode(const PassArgs* d_args, const PassDyn& dy, uint64_t ntiles, hipStream_t st) {
  constexpr uint64_t per_block = kThreads / 64;
  const uint64_t maxb = 1ull << 30;
  PassDyn d = dy;
  for (uint64_t t0 = 0; t0 < ntiles; t0 += maxb * per_block) {
    const uint64_t cnt = std::min<uint64_t>(ntiles - t0, maxb * per_block);
    d.tile0 = dy.tile0 + t0;
    hipLaunchKernelGGL((wave_kernel<IN, OUT>), dim3((unsigned)((cnt + per_block - 1) / per_block)), dim3(kThreads), 0,
                       st, d_args, d, dy.tile0 + ntiles);
  }
  return hipGetLastError();
}

// ntiles counts wave tiles (2^10 amplitudes each) of this launch
hipError_t launch_wave(int in_mode, int out_mode, const PassArgs* d_args, const PassDyn& dy, uint64_t ntiles,
                       hipStream_t st) {
  if (in_mode == IN_RAW && out_mode == OUT_RAW) return launch_wave_mode<IN_RAW, OUT_RAW>(d_args, dy, ntiles, st);
  if (in_mode == IN_DEC && out_mode == OUT_RAW) return launch_wave_mode<IN_DEC, OUT_RAW>(d_args, dy, ntiles, st);
  if (in_mode == IN_RAW && out_mode == OUT_ENC) return launch_wave_mode<IN_RAW, OUT_ENC>(d_args, dy, ntiles, st);
  return launch_wave_mode<IN_DEC, OUT_ENC>(d_args, dy, ntiles, st);
}

}  // namespace qsim
