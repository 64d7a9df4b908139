// Device helpers shared by the LDS-tile pass kernel (qsim_kernels.hip) and the
// wave-resident pass kernel (qsim_wave.hip): bit insertion, tile maps, wave64
// DPP reductions, and the chunk codec.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include <stdlib.h>
#include "qsim_internal.h"

// Address-space qualifiers: pointers read out of argument structs are
// generic to the compiler, which then emits flat memory instructions; casting
// to global (1) or constant (4) memory gives global / scalar loads.
#define QG __attribute__((address_space(1)))
#define QC __attribute__((address_space(4)))

namespace qsim {

// ------------------------------------------------------- launch shapes (host)
// Work-groups per dispatch: launches are split at kMaxLaunchBlocks groups
// (the dispatch packet's 32-bit work-item count).  QOMPSIM_MAX_LAUNCH_GROUPS
// lowers the split point, so that multi-dispatch launches can be exercised at
// small sizes.
inline uint64_t max_launch_groups() {
  const char* e = std::getenv("QOMPSIM_MAX_LAUNCH_GROUPS");
  const long long v = e ? std::atoll(e) : 0;
  return (v > 0 && (uint64_t)v < kMaxLaunchBlocks) ? (uint64_t)v : kMaxLaunchBlocks;
}
// Work-groups of a persistent grid (each wave walks its tiles with the
// grid's stride); QOMPSIM_PERSISTENT_GROUPS overrides the default.
inline uint64_t persistent_groups(uint64_t dflt) {
  const char* e = std::getenv("QOMPSIM_PERSISTENT_GROUPS");
  const long long v = e ? std::atoll(e) : 0;
  return (v > 0 && (uint64_t)v < kMaxLaunchBlocks) ? (uint64_t)v : dflt;
}

// Waves of a persistent wave-kernel launch over ntiles tiles (four waves per
// work-group, at most persistent_groups(16384) work-groups: 65536 waves, one
// sub-arena each in the fused DEC->ENC form).  QFT-32 at 32 qubits: 609 ms
// per step at 8192 work-groups, 601 ms at 16384.
inline uint64_t persistent_waves(uint64_t ntiles) {
  const uint64_t cap = persistent_groups(16384);
  const uint64_t groups = (ntiles + 3) / 4 < cap ? (ntiles + 3) / 4 : cap;
  return groups * 4;
}

// ---------------------------------------------------------------- helpers
__device__ __forceinline__ uint32_t ins0(uint32_t x, uint32_t pos) {
  uint32_t lo = x & ((1u << pos) - 1u);
  return lo | ((x >> pos) << (pos + 1));
}
__device__ __forceinline__ uint64_t ins0_64(uint64_t x, uint32_t pos) {
  uint64_t lo = x & ((1ull << pos) - 1ull);
  return lo | ((x >> pos) << (pos + 1));
}

// Allocation piece of a tile in the region store: the tile index folded so
// that pieces mix tiles of every low-bit pattern (payload sizes can follow
// the tile index's low bits, which select fixed amplitude positions).
__device__ __forceinline__ uint32_t region_piece(uint64_t t, uint32_t mask) {
  return (uint32_t)(t ^ (t >> 6) ^ (t >> 12) ^ (t >> 18) ^ (t >> 24)) & mask;
}

template <class EPtr>
struct TileMapT {
  uint32_t t, c, nE;
  EPtr E;
  __device__ __forceinline__ uint64_t base(uint64_t tau) const {
    uint64_t x = tau << c;
    for (uint32_t i = 0; i < nE; ++i) x = ins0_64(x, E[i]);
    return x;
  }
  __device__ __forceinline__ uint64_t off(uint32_t p) const {
    uint64_t o = p & ((1u << c) - 1u);
    for (uint32_t i = 0; i < nE; ++i) o |= (uint64_t)((p >> (c + i)) & 1u) << E[i];
    return o;
  }
};
using TileMapDev = TileMapT<const uint32_t*>;
using TileMapConst = TileMapT<const QC uint32_t*>;

// (cos a, sin a) for |a| < 2^20: Cody-Waite reduction modulo pi/2 (pi/2 split
// into three parts; the first two carry 33 significant bits, so k*P1 and k*P2
// are exact for |k| < 2^20) and the Taylor series of sin and cos on
// [-pi/4, pi/4] through degree 17 and 18 (truncation below 1e-17).  The
// coefficients are 1/n! rounded to double.  Phase angles reaching the kernels
// are sums of terms in [-pi, pi], far inside the range.
__device__ __forceinline__ double2 cis_reduced(double a) {
  const double k = rint(a * 0.6366197723675814);  // 2/pi
  double r = fma(-k, 1.5707963267341256, a);
  r = fma(-k, 6.077100506303966e-11, r);
  r = fma(-k, 2.0222662487959506e-21, r);
  const double r2 = r * r;
  double q = 2.8114572543455206e-15;        //  1/17!
  q = fma(q, r2, -7.647163731819816e-13);   // -1/15!
  q = fma(q, r2, 1.6059043836821613e-10);   //  1/13!
  q = fma(q, r2, -2.505210838544172e-08);   // -1/11!
  q = fma(q, r2, 2.7557319223985893e-06);   //  1/9!
  q = fma(q, r2, -0.0001984126984126984);   // -1/7!
  q = fma(q, r2, 0.008333333333333333);     //  1/5!
  q = fma(q, r2, -0.16666666666666666);     // -1/3!
  const double sr = fma(q * r2, r, r);
  double p = -1.5619206968586225e-16;       // -1/18!
  p = fma(p, r2, 4.779477332387385e-14);    //  1/16!
  p = fma(p, r2, -1.1470745597729725e-11);  // -1/14!
  p = fma(p, r2, 2.08767569878681e-09);     //  1/12!
  p = fma(p, r2, -2.755731922398589e-07);   // -1/10!
  p = fma(p, r2, 2.48015873015873e-05);     //  1/8!
  p = fma(p, r2, -0.001388888888888889);    // -1/6!
  p = fma(p, r2, 0.041666666666666664);     //  1/4!
  p = fma(p, r2, -0.5);                     // -1/2!
  const double cr = fma(p, r2, 1.0);
  const int n = (int)k & 3;
  const double c0 = (n & 1) ? sr : cr, s0 = (n & 1) ? cr : sr;
  return make_double2(((n + 1) & 2) ? -c0 : c0, (n & 2) ? -s0 : s0);
}

// wave64 reductions through DPP (quad perms, row mirrors, row broadcasts);
// result is read from lane 63.  All 64 lanes must be active.
__device__ __forceinline__ int wave_min_i32(int v) {
  const int id = 0x7fffffff;
  v = min(v, __builtin_amdgcn_update_dpp(id, v, 0xb1, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(id, v, 0x4e, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(id, v, 0x141, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(id, v, 0x140, 0xf, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(id, v, 0x142, 0xa, 0xf, false));
  v = min(v, __builtin_amdgcn_update_dpp(id, v, 0x143, 0xc, 0xf, false));
  return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int wave_max_i32(int v) {
  const int id = (int)0x80000000;
  v = max(v, __builtin_amdgcn_update_dpp(id, v, 0xb1, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(id, v, 0x4e, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(id, v, 0x141, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(id, v, 0x140, 0xf, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(id, v, 0x142, 0xa, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(id, v, 0x143, 0xc, 0xf, false));
  return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ uint64_t ballot64(bool p) { return __ballot(p); }

// ------------------------------------------------------------------ codec
struct CodecParams {
  double eps_rel, eps_abs;
  int guard;
  int e_abs;      // ilogb(2*eps_abs), or INT_MIN when eps_abs == 0
  double p_abs;   // 2^e_abs
};

// Per-lane exponent candidate; INT_MAX for zero lanes, INT_MIN requests the
// verbatim mode (non-finite value or a zero error budget).
// A non-finite x gives a non-finite y, so one class test (y positive normal
// or subnormal) covers every verbatim case.
__device__ __forceinline__ int lane_exponent(double x, const CodecParams& cp) {
  const double y = 2.0 * fma(cp.eps_rel, fabs(x), cp.eps_abs);
  const int ey = __builtin_amdgcn_frexp_exp(y) - 1;  // == ilogb(y)
  const bool ok = __builtin_amdgcn_class(y, 0x180);  // +normal | +subnormal
  return x == 0.0 ? 0x7fffffff : (ok ? ey : (int)0x80000000);
}

// Quantise one chunk (wave-uniform call).  Returns w (0, 1..52, or 64) and e.
// e is the minimum over nonzero lanes of ilogb(2*fma(eps_rel,|x|,eps_abs))
// minus the guard bits.  Every lane has y >= 2*eps_abs, so when one nonzero
// lane has y < 2^(e_abs+1) the minimum is e_abs itself and the wave-wide
// reduction is skipped (same result).
__device__ __forceinline__ void chunk_plan(double x, const CodecParams& cp, uint32_t& w, int& e) {
  const bool nz = x != 0.0;
  if (ballot64(nz) == 0) { w = 0; e = 0; return; }
  int emin;
  const bool small = nz && (fma(cp.eps_rel, fabs(x), cp.eps_abs) < cp.p_abs);
  if (cp.e_abs != (int)0x80000000 && ballot64(small) != 0) emin = cp.e_abs;
  else emin = wave_min_i32(lane_exponent(x, cp));
  if (emin == (int)0x80000000 || emin < -1000) { w = 64; e = 0; return; }
  e = emin - cp.guard;
  const double s = ldexp(x, -e);
  const bool bad = !(fabs(s) < 2251799813685248.0);  // 2^51 (also NaN / inf)
  if (ballot64(bad) != 0) { w = 64; e = 0; return; }
  const double k = rint(s);
  if (ballot64(!(fabs(k) < 1073741824.0)) == 0) {    // |k| < 2^30: 32-bit codes
    const int ki = (int)k;
    const uint32_t u = ((uint32_t)ki << 1) ^ (uint32_t)(ki >> 31);
    const int bl = (u == 0) ? 0 : 32 - __clz((int)u);
    w = (uint32_t)wave_max_i32(bl);
    return;
  }
  const double uz = (k >= 0.0) ? 2.0 * k : -2.0 * k - 1.0;
  const uint64_t u = (uint64_t)uz;
  const int bl = (u == 0) ? 0 : 64 - __clzll(u);
  w = (uint32_t)wave_max_i32(bl);
}

__device__ __forceinline__ uint64_t chunk_code(double x, uint32_t w, int e) {
  if (w == 64) return (uint64_t)__double_as_longlong(x);
  if (w == 0) return 0;
  const double k = rint(ldexp(x, -e));
  const double uz = (k >= 0.0) ? 2.0 * k : -2.0 * k - 1.0;
  return (uint64_t)uz;
}

// zig-zag code of rint(x * 2^-e) for chunks whose codes fit 32 bits
__device__ __forceinline__ uint32_t chunk_code32(double x, int e) {
  const int ki = (int)rint(ldexp(x, -e));
  return ((uint32_t)ki << 1) ^ (uint32_t)(ki >> 31);
}

__device__ __forceinline__ double decode_lane(uint64_t d, const uint64_t* __restrict__ arena, int lane) {
  const uint32_t w = desc_w(d);
  if (w == 0) return 0.0;
  const uint64_t* P = arena + desc_off(d);
  if (w == 64) return __longlong_as_double((long long)P[lane]);
  if (w <= 32) {
    const uint32_t* P32 = reinterpret_cast<const uint32_t*>(P);
    const uint32_t bit = (uint32_t)lane * w;
    const uint32_t q = bit >> 5, s = bit & 31u;
    const uint32_t lo = P32[q];
    const uint32_t hi = (s + w > 32u) ? P32[q + 1] : 0u;
    uint32_t v = (uint32_t)((((uint64_t)hi << 32) | lo) >> s);
    if (w < 32) v &= (1u << w) - 1u;
    const int k = (int)(v >> 1) ^ -(int)(v & 1u);
    return ldexp((double)k, desc_e(d));
  }
  const uint32_t bit = (uint32_t)lane * w;
  const uint32_t q = bit >> 6, s = bit & 63u;
  uint64_t v = P[q] >> s;
  if (s + w > 64u) v |= P[q + 1] << (64u - s);
  v &= (1ull << w) - 1ull;
  const int64_t k = (int64_t)(v >> 1) ^ -(int64_t)(v & 1ull);
  return ldexp((double)k, desc_e(d));
}


__device__ __forceinline__ void cmad(double& re, double& im, double mr, double mi, double vr, double vi) {
  re = fma(mr, vr, re);
  re = fma(-mi, vi, re);
  im = fma(mr, vi, im);
  im = fma(mi, vr, im);
}

}  // namespace qsim
