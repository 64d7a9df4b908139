// HIP kernels for gfx950 (CDNA4, wave64).
//
// pass_kernel<IN, OUT, TB>: one work-group owns one tile of 2^TB amplitudes of
// a group-local state vector.  It loads the tile (raw complex128 from HBM, or
// decodes codec chunks of the compressed block store), applies every op of the
// pass in LDS, and stores it back (raw, or re-encoded into the store with the
// all-zero check of blocks that were null at stage start).
//
// Codec (per 64-double chunk, one double per lane):
//   e   = min over nonzero lanes of ilogb(2*fma(eps_rel,|x|,eps_abs)) - guard
//   k   = rint(x * 2^-e)       |x - k*2^e| <= 2^(e-1) <= eps_abs + eps_rel*|x|
//   u   = zig-zag(k), w = max bit length over the wave (wave64 DPP reduction)
//   out = lane-major bit packing, chunk offsets by a work-group prefix sum and
//         one atomic per tile on one of 64 sub-arena cursors.
// Chunks whose codes would need more than 52 bits are stored verbatim (w=64).
#include <hip/hip_runtime.h>
#include <mutex>
#include <stdint.h>
#include <math.h>
#include "qsim_internal.h"
#include "qsim_device.h"

namespace qsim {

// ------------------------------------------------------------ pass kernel


template <int K, int TB>
__device__ void dense_op(const Op& op, double2* tile, uint64_t base, const double* __restrict__ coefs) {
  constexpr uint32_t NG = 1u << (TB - K);
  constexpr int D = 1 << K;
  uint32_t ts[K], tp[K];
#pragma unroll
  for (int j = 0; j < K; ++j) { ts[j] = op.tsorted[j]; tp[j] = op.tpos[j]; }
  const uint32_t nsel = op.nsel;
  const uint64_t ident = op.ident_mask;
  // selector contribution of base bits (uniform over the tile)
  uint32_t sel_base = 0, sel_tile_mask = 0;
  for (uint32_t s = 0; s < nsel; ++s) {
    if (op.sel_tile[s]) sel_tile_mask |= 1u << s;
    else sel_base |= (uint32_t)((base >> op.sel_pos[s]) & 1ull) << s;
  }
  for (uint32_t gi = threadIdx.x; gi < NG; gi += kThreads) {
    uint32_t i0 = gi;
#pragma unroll
    for (int j = 0; j < K; ++j) i0 = ins0(i0, ts[j]);
    uint32_t sel = sel_base;
    if (sel_tile_mask) {
      for (uint32_t s = 0; s < nsel; ++s)
        if ((sel_tile_mask >> s) & 1u) sel |= ((i0 >> op.sel_pos[s]) & 1u) << s;
    }
    if ((ident >> sel) & 1ull) continue;
    const double* M = coefs + 2ull * (op.mat_off + (uint64_t)sel * (uint64_t)(D * D));
    uint32_t idx[D];
    double vr[D], vi[D];
#pragma unroll
    for (int cc = 0; cc < D; ++cc) {
      uint32_t o = i0;
#pragma unroll
      for (int j = 0; j < K; ++j) if ((cc >> j) & 1) o |= 1u << tp[j];
      idx[cc] = o;
      const double2 v = tile[o];
      vr[cc] = v.x; vi[cc] = v.y;
    }
#pragma unroll
    for (int r = 0; r < D; ++r) {
      double re = 0.0, im = 0.0;
#pragma unroll
      for (int cc = 0; cc < D; ++cc) cmad(re, im, M[2 * (r * D + cc)], M[2 * (r * D + cc) + 1], vr[cc], vi[cc]);
      tile[idx[r]] = make_double2(re, im);
    }
  }
}

// Phase table entries per tile: 64 for the low six tile bits, up to 128 for
// the rest.
constexpr int kPhaseTab = 192;

// A run of diagonal gates.  Phase terms whose tile condition involves only
// the low six tile bits (or none) are folded into tlo[64], those involving only
// the high tile bits into thi[]; both tables are built once per tile (one
// sincos per entry), so an amplitude costs two LDS reads and two complex
// multiplies.  Terms spanning both halves are summed per amplitude.
template <int TB>
__device__ void elem_op(const Op& op, double2* tile, uint64_t base, const PhaseTerm* __restrict__ terms,
                        const DiagTable* __restrict__ diags, const double* __restrict__ coefs, double2* ptab) {
  constexpr uint32_t NT = 1u << TB;
  constexpr int EPT = (NT + kThreads - 1) / kThreads;
  constexpr uint32_t LOB = TB < 6 ? TB : 6;
  constexpr uint32_t NLO = 1u << LOB, NHI = 1u << (TB - LOB);
  constexpr uint32_t LOM = NLO - 1u;
  const uint32_t nterms = op.nterms;
  if (nterms) {
    const uint32_t tid = threadIdx.x;
    if (tid < NLO + NHI) {
      const bool lo = tid < NLO;
      const uint32_t idx = lo ? tid : tid - NLO;
      double ang = 0.0;
      for (uint32_t i = 0; i < nterms; ++i) {
        const PhaseTerm& T = terms[op.term_off + i];
        if ((base & T.bmask) != T.bval) continue;
        const uint32_t tl = T.tmask & LOM, th = T.tmask >> LOB;
        if (tl && th) continue;
        if (lo) {
          if (!th) ang += ((idx & tl) == (T.tval & LOM)) ? T.angle : 0.0;
        } else {
          if (th) ang += ((idx & th) == (T.tval >> LOB)) ? T.angle : 0.0;
        }
      }
      double sn, cs;
      sincos(ang, &sn, &cs);
      ptab[lo ? idx : 64 + idx] = make_double2(cs, sn);
    }
    __syncthreads();
  }
  double ang[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) ang[e] = 0.0;
  bool cross = false;
  for (uint32_t i = 0; i < nterms; ++i) {
    const PhaseTerm& T = terms[op.term_off + i];
    if (!((T.tmask & LOM) && (T.tmask >> LOB))) continue;
    if ((base & T.bmask) != T.bval) continue;
    cross = true;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const uint32_t p = threadIdx.x + e * kThreads;
      ang[e] += ((p & T.tmask) == T.tval) ? T.angle : 0.0;
    }
  }
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const uint32_t p = threadIdx.x + e * kThreads;
    if (NT < kThreads && p >= NT) break;
    double2 v = tile[p];
    if (nterms) {
      const double2 a = ptab[p & LOM];
      const double2 b = ptab[64 + (p >> LOB)];
      const double2 ab = make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
      v = make_double2(v.x * ab.x - v.y * ab.y, v.x * ab.y + v.y * ab.x);
    }
    if (cross) {
      double sn, cs;
      sincos(ang[e], &sn, &cs);
      v = make_double2(v.x * cs - v.y * sn, v.x * sn + v.y * cs);
    }
    for (uint32_t d = 0; d < op.ndiag; ++d) {
      const DiagTable& Dt = diags[op.diag_off + d];
      uint32_t idx = 0;
      for (uint32_t j = 0; j < Dt.m; ++j) {
        const uint32_t bit = Dt.tile[j] ? ((p >> Dt.pos[j]) & 1u) : (uint32_t)((base >> Dt.pos[j]) & 1ull);
        idx |= bit << j;
      }
      const double cr = coefs[2ull * (Dt.coef_off + idx)], ci = coefs[2ull * (Dt.coef_off + idx) + 1];
      v = make_double2(v.x * cr - v.y * ci, v.x * ci + v.y * cr);
    }
    tile[p] = v;
  }
}

// BIG instantiations carry the 4- and 5-target dense ops; the common kernel
// stops at 3 targets so its register budget stays small.
template <int TB, bool BIG>
__device__ void apply_ops(const PassArgs& a, double2* tile, uint64_t base, double2* ptab) {
  for (uint32_t o = a.op_begin; o < a.op_end; ++o) {
    const Op& op = a.ops[o];
    if (op.kind == OP_DENSE) {
      switch (op.k) {
        case 1: if constexpr (TB >= 1) dense_op<1, TB>(op, tile, base, a.coefs); break;
        case 2: if constexpr (TB >= 2) dense_op<2, TB>(op, tile, base, a.coefs); break;
        case 3: if constexpr (TB >= 3) dense_op<3, TB>(op, tile, base, a.coefs); break;
        case 4: if constexpr (BIG && TB >= 4) dense_op<4, TB>(op, tile, base, a.coefs); break;
        case 5: if constexpr (BIG && TB >= 5) dense_op<5, TB>(op, tile, base, a.coefs); break;
        default: break;
      }
    } else {
      elem_op<TB>(op, tile, base, a.terms, a.diags, a.coefs, ptab);
    }
    __syncthreads();
  }
}

// LDS layout beyond the tile (codec modes only)
struct CodecScratch {
  uint64_t desc[1u << (kMaxTileBits - kChunkAmpsLog2)];
  uint32_t bid[1u << (kMaxTileBits - kChunkAmpsLog2)];
  uint32_t ci[1u << (kMaxTileBits - kChunkAmpsLog2)];
  uint32_t w[1u << (kMaxTileBits - kChunkAmpsLog2)];
  int32_t e[1u << (kMaxTileBits - kChunkAmpsLog2)];
  uint32_t pre[1u << (kMaxTileBits - kChunkAmpsLog2)];
  uint32_t wave_tot[kThreads / 64];
  unsigned long long tile_base;
  uint32_t skip;
  uint64_t stage[kThreads / 64][64];
};

template <int IN, int OUT, int TB, bool BIG>
__global__ __launch_bounds__(kThreads) void pass_kernel(const PassArgs* __restrict__ pa, PassDyn dy) {
  extern __shared__ __align__(16) unsigned char smem[];
  double2* tile = reinterpret_cast<double2*>(smem);
  constexpr uint32_t NT = 1u << TB;
  const PassArgs& a = *pa;
  const TileMapDev tm{a.t, a.c, a.nE, a.E};
  const uint32_t lbits = a.L;
  const uint64_t tau_all = dy.tile0 + blockIdx.x;
  const uint64_t slot = tau_all >> (lbits - TB);
  const uint64_t tau = tau_all & ((1ull << (lbits - TB)) - 1ull);
  const uint64_t base = tm.base(tau);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double2* ptab = reinterpret_cast<double2*>(smem + sizeof(double2) * NT);
  CodecScratch* cs = reinterpret_cast<CodecScratch*>(smem + sizeof(double2) * (NT + kPhaseTab));
  constexpr uint32_t NCH = (TB >= kChunkAmpsLog2) ? (NT >> kChunkAmpsLog2) : 1u;
  const uint64_t bmask = (1ull << a.B) - 1ull;

  // ---- chunk metadata (codec modes): one thread per chunk
  if constexpr (IN == IN_DEC || OUT == OUT_ENC) {
    for (uint32_t j = threadIdx.x; j < NCH; j += kThreads) {
      const uint64_t loc = base | tm.off(j << kChunkAmpsLog2);
      const uint64_t bpos = (slot << a.khi) + (loc >> a.B);
      const uint32_t bid = dy.bmap[bpos];
      const uint32_t ci = (uint32_t)((loc & bmask) >> kChunkAmpsLog2);
      cs->bid[j] = bid;
      cs->ci[j] = ci;
      if constexpr (IN == IN_DEC) {
        cs->desc[j] = dy.null_in[bid] ? 0ull : dy.desc_in[(dy.desc_staged ? bpos : (uint64_t)bid) * a.cpb + ci];
      }
    }
    __syncthreads();
  }

  // ---- load
  if constexpr (IN == IN_RAW) {
    const double2* src = reinterpret_cast<const double2*>(dy.raw) + (slot << lbits) + base;
#pragma unroll
    for (uint32_t e = 0; e < (NT + kThreads - 1) / kThreads; ++e) {
      const uint32_t p = threadIdx.x + e * kThreads;
      if (NT >= kThreads || p < NT) tile[p] = src[tm.off(p)];
    }
  } else {
    double* td = reinterpret_cast<double*>(tile);
    constexpr int U = 4;
    for (uint32_t j0 = wave; j0 < NCH; j0 += U * (kThreads / 64)) {
      double xv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t j = j0 + u * (kThreads / 64);
        xv[u] = (j < NCH) ? decode_lane(cs->desc[j], dy.arena_in, lane) : 0.0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t j = j0 + u * (kThreads / 64);
        if (j < NCH) td[(j << (kChunkAmpsLog2 + 1)) + lane] = xv[u];
      }
    }
  }
  __syncthreads();

  apply_ops<TB, BIG>(a, tile, base, ptab);

  // ---- store
  if constexpr (OUT == OUT_RAW) {
    double2* dst = reinterpret_cast<double2*>(dy.raw) + (slot << lbits) + base;
#pragma unroll
    for (uint32_t e = 0; e < (NT + kThreads - 1) / kThreads; ++e) {
      const uint32_t p = threadIdx.x + e * kThreads;
      if (NT >= kThreads || p < NT) dst[tm.off(p)] = tile[p];
    }
  } else {
    const double* td = reinterpret_cast<const double*>(tile);
    const CodecParams cp{a.eps_rel, a.eps_abs, a.guard, a.e_abs, a.p_abs};
    // phase 1: widths, exponents, zero check of blocks null at stage start
    for (uint32_t j = wave; j < NCH; j += kThreads / 64) {
      const double x = td[(j << (kChunkAmpsLog2 + 1)) + lane];
      const uint32_t bid = cs->bid[j];
      if (dy.null_in[bid]) {
        const bool nz = (a.zero_mode == ZERO_EXACT) ? (x != 0.0) : (fabs(x) > a.zero_thr);
        if (ballot64(nz) != 0 && lane == 0) dy.nonzero[bid] = 1;
      }
      uint32_t w; int e;
      chunk_plan(x, cp, w, e);
      if (lane == 0) { cs->w[j] = w; cs->e[j] = e; }
    }
    __syncthreads();
    // exclusive prefix of chunk sizes (in 8-byte words)
    {
      uint32_t v = (threadIdx.x < NCH) ? cs->w[threadIdx.x] : 0u;
      uint32_t incl = v;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(incl, d, 64);
        if (lane >= d) incl += y;
      }
      if (lane == 63) cs->wave_tot[wave] = incl;
      __syncthreads();
      uint32_t woff = 0, tot = 0;
#pragma unroll
      for (int q = 0; q < kThreads / 64; ++q) {
        if (q < wave) woff += cs->wave_tot[q];
        tot += cs->wave_tot[q];
      }
      if (threadIdx.x < NCH) cs->pre[threadIdx.x] = woff + incl - v;
      if (threadIdx.x == 0) {
        const uint32_t sub = dy.slot_region
                                 ? 2u * (dy.slot_region[slot] * (dy.sub_mask + 1u) + region_piece(blockIdx.x, dy.sub_mask))
                                 : (blockIdx.x & dy.sub_mask);
        unsigned long long off = 0;
        uint32_t skip = 0;
        if (tot > 0) {
          off = atomicAdd(&dy.cursors[sub * 16], (unsigned long long)tot);
          if (off + tot > dy.sub_base[sub + 1] - dy.sub_base[sub]) { skip = 1; atomicOr(dy.overflow, 1u); }
        }
        cs->tile_base = off + dy.sub_base[sub];
        cs->skip = skip;
      }
      cs->stage[wave][lane] = 0ull;
      __syncthreads();
    }
    // phase 2: pack and write
    const unsigned long long tbase = cs->tile_base;
    const uint32_t skip = cs->skip;
    for (uint32_t j = wave; j < NCH; j += kThreads / 64) {
      const uint32_t w = cs->w[j];
      const int e = cs->e[j];
      const uint64_t off = tbase + cs->pre[j];
      if (lane == 0) dy.desc_out[(uint64_t)cs->bid[j] * a.cpb + cs->ci[j]] = desc_pack(off, w, e);
      if (w == 0 || skip) continue;
      const double x = td[(j << (kChunkAmpsLog2 + 1)) + lane];
      if (w <= 32) {
        uint32_t* st32 = reinterpret_cast<uint32_t*>(cs->stage[wave]);
        const uint32_t u = chunk_code32(x, e);
        const uint32_t bit = (uint32_t)lane * w;
        const uint32_t q = bit >> 5, s = bit & 31u;
        atomicOr(&st32[q], u << s);
        if (s + w > 32u) atomicOr(&st32[q + 1], u >> (32u - s));
        __builtin_amdgcn_wave_barrier();
        uint32_t* out32 = reinterpret_cast<uint32_t*>(dy.arena_out + off);
        if ((uint32_t)lane < 2u * w) {
          out32[lane] = st32[lane];
          st32[lane] = 0u;
        }
        __builtin_amdgcn_wave_barrier();
        continue;
      }
      const uint64_t u = chunk_code(x, w, e);
      const uint32_t bit = (uint32_t)lane * w;
      const uint32_t q = bit >> 6, s = bit & 63u;
      atomicOr(&cs->stage[wave][q], (unsigned long long)(u << s));
      if (s + w > 64u) atomicOr(&cs->stage[wave][q + 1], (unsigned long long)(u >> (64u - s)));
      __builtin_amdgcn_wave_barrier();
      if ((uint32_t)lane < w) {
        dy.arena_out[off + lane] = cs->stage[wave][lane];
        cs->stage[wave][lane] = 0ull;
      }
      __builtin_amdgcn_wave_barrier();
    }
  }
}

// ------------------------------------------------------- state utilities
__global__ void fill_basis_kernel(double2* v, uint64_t n, uint64_t idx) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t j = i; j < n; j += stride) v[j] = make_double2(j == idx ? 1.0 : 0.0, 0.0);
}

__global__ void fill_u8_kernel(uint8_t* p, uint64_t n, uint8_t val) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t j = i; j < n; j += stride) p[j] = val;
}

__global__ void fill_u64_kernel(unsigned long long* p, uint64_t n, unsigned long long val) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t j = i; j < n; j += stride) p[j] = val;
}

__global__ void scale_kernel(double2* v, uint64_t n, double s) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t j = i; j < n; j += stride) { double2 x = v[j]; v[j] = make_double2(x.x * s, x.y * s); }
}

// null <- null && !nonzero ; nonzero cleared
__global__ void commit_null_kernel(uint8_t* null_flags, uint8_t* nonzero, uint64_t nb) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nb) {
    if (nonzero[i]) null_flags[i] = 0;
    nonzero[i] = 0;
  }
}

// ------------------------------------------------ compensated reductions
struct DD { double s, c; };
__device__ __forceinline__ void dd_add(double& s, double& c, double x) {
  const double t = s + x;
  const double bp = t - s;
  c += (s - (t - bp)) + (x - bp);
  s = t;
}
__device__ __forceinline__ void dd_merge(double& s, double& c, double s2, double c2) {
  dd_add(s, c, s2);
  c += c2;
}

// Block reduce of NV compensated accumulators; writes partials[blockIdx][NV][2].
template <int NV>
__device__ void block_reduce_dd(double (&s)[NV], double (&c)[NV], double* partials) {
  __shared__ double sh_s[NV][kThreads / 64];
  __shared__ double sh_c[NV][kThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    for (int d = 32; d >= 1; d >>= 1) {
      const double s2 = __shfl_xor(s[v], d, 64);
      const double c2 = __shfl_xor(c[v], d, 64);
      dd_merge(s[v], c[v], s2, c2);
    }
    if (lane == 0) { sh_s[v][wave] = s[v]; sh_c[v][wave] = c[v]; }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      double ss = sh_s[v][0], cc = sh_c[v][0];
      for (int q = 1; q < kThreads / 64; ++q) dd_merge(ss, cc, sh_s[v][q], sh_c[v][q]);
      partials[(blockIdx.x * NV + v) * 2] = ss;
      partials[(blockIdx.x * NV + v) * 2 + 1] = cc;
    }
  }
}

// mode 0: sum |a|^2 ; mode 1: sum conj(a)*b (re, im)
template <int MODE>
__global__ __launch_bounds__(kThreads) void sv_reduce_kernel(const double2* __restrict__ a, const double2* __restrict__ b,
                                                             uint64_t n, double* partials) {
  constexpr int NV = MODE == 0 ? 1 : 2;
  double s[NV], c[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) { s[v] = 0.0; c[v] = 0.0; }
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    const double2 x = a[i];
    if constexpr (MODE == 0) {
      dd_add(s[0], c[0], fma(x.x, x.x, x.y * x.y));
    } else {
      const double2 y = b[i];
      dd_add(s[0], c[0], fma(x.x, y.x, x.y * y.y));
      dd_add(s[1], c[1], fma(x.x, y.y, -x.y * y.x));
    }
  }
  block_reduce_dd<NV>(s, c, partials);
}

// <a| P |a> accumulation for one Pauli string: sum_i sign(i) conj(a[i^f]) a[i]
// (the i^{nY} factor is applied by the host).
__global__ __launch_bounds__(kThreads) void sv_pauli_kernel(const double2* __restrict__ a, uint64_t n, uint64_t fmask,
                                                            uint64_t zmask, double* partials) {
  double s[2] = {0.0, 0.0}, c[2] = {0.0, 0.0};
  const uint64_t stride = (uint64_t)gridDim.x * kThreads;
  for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    const double2 x = a[i];
    const double2 y = a[i ^ fmask];
    double re = fma(y.x, x.x, y.y * x.y);
    double im = fma(y.x, x.y, -y.y * x.x);
    if (__popcll(i & zmask) & 1) { re = -re; im = -im; }
    dd_add(s[0], c[0], re);
    dd_add(s[1], c[1], im);
  }
  block_reduce_dd<2>(s, c, partials);
}

// out[j] (+)= coef * P a[j]; P|i> = i^nY (-1)^{pc(i & zy)} |i ^ f>, so
// (P a)[j] = i^nY (-1)^{pc((j^f) & zy)} a[j^f]; the complex factor coef*i^nY
// is passed in (cr, ci).
__global__ void sv_apply_pauli_kernel(const double2* __restrict__ a, double2* __restrict__ out, uint64_t n,
                                      uint64_t fmask, uint64_t zmask, double cr, double ci, int accumulate) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
    const uint64_t i = j ^ fmask;
    double2 x = a[i];
    double sg = (__popcll(i & zmask) & 1) ? -1.0 : 1.0;
    const double re = sg * (cr * x.x - ci * x.y);
    const double im = sg * (cr * x.y + ci * x.x);
    if (accumulate) { double2 o = out[j]; out[j] = make_double2(o.x + re, o.y + im); }
    else out[j] = make_double2(re, im);
  }
}

// payload + descriptor bytes of every block (0 for null blocks)
__global__ __launch_bounds__(256) void block_bytes_kernel(const uint64_t* __restrict__ desc, const uint8_t* __restrict__ nullf,
                                                          uint64_t nb, uint64_t cpb, unsigned long long* out) {
  __shared__ unsigned long long sh[256];
  for (uint64_t b = blockIdx.x; b < nb; b += gridDim.x) {
    unsigned long long s = 0;
    if (!nullf[b])
      for (uint64_t c = threadIdx.x; c < cpb; c += 256) s += desc_w(desc[b * cpb + c]);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (unsigned d = 128; d >= 1; d >>= 1) {
      if (threadIdx.x < d) sh[threadIdx.x] += sh[threadIdx.x + d];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[b] = nullf[b] ? 0ull : 8ull * (sh[0] + cpb);
    __syncthreads();
  }
}

// sum and maximum of the allocation cursors (stride 16 words)
// Sum and maximum of the n allocation cursors (one per 128-byte line):
// kCursorBlocks work-groups each reduce a slice into a partial pair
// (part[2b], part[2b + 1]); one more work-group combines the pairs into out.
constexpr unsigned kCursorBlocks = 64;
__global__ __launch_bounds__(256) void cursor_stats_kernel(const unsigned long long* __restrict__ cur, uint32_t n,
                                                           unsigned long long* __restrict__ part) {
  __shared__ unsigned long long ss[256], sm[256];
  unsigned long long s = 0, m = 0;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += kCursorBlocks * 256) {
    const unsigned long long v = cur[(uint64_t)i * 16];
    s += v;
    m = v > m ? v : m;
  }
  ss[threadIdx.x] = s;
  sm[threadIdx.x] = m;
  __syncthreads();
  for (unsigned d = 128; d >= 1; d >>= 1) {
    if (threadIdx.x < d) {
      ss[threadIdx.x] += ss[threadIdx.x + d];
      sm[threadIdx.x] = sm[threadIdx.x] > sm[threadIdx.x + d] ? sm[threadIdx.x] : sm[threadIdx.x + d];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = ss[0];
    part[2 * blockIdx.x + 1] = sm[0];
  }
}
__global__ __launch_bounds__(64) void cursor_stats_final_kernel(const unsigned long long* __restrict__ part,
                                                                unsigned long long* __restrict__ out) {
  if (threadIdx.x != 0) return;
  unsigned long long s = 0, m = 0;
  for (unsigned b = 0; b < kCursorBlocks; ++b) {
    s += part[2 * b];
    m = part[2 * b + 1] > m ? part[2 * b + 1] : m;
  }
  out[0] = s;
  out[1] = m;
}

// Descriptors of a launch's blocks in bmap order (the region store encodes
// its output descriptors in place, so a stage reads its input from this
// copy): out[i cpb + c] = desc[bmap[i] cpb + c].
__global__ __launch_bounds__(256) void gather_desc_kernel(const uint64_t* __restrict__ desc,
                                                          const uint32_t* __restrict__ bmap, uint64_t nblk,
                                                          uint64_t cpb, uint64_t* __restrict__ out) {
  const uint64_t n = nblk * cpb;
  for (uint64_t x = (uint64_t)blockIdx.x * 256 + threadIdx.x; x < n; x += (uint64_t)gridDim.x * 256)
    out[x] = desc[(uint64_t)bmap[x / cpb] * cpb + x % cpb];
}

// ------------------------------------------------ adjoint reverse sweep
// Reverse sweep of the adjoint method (ParametricQuantumCircuit::
// backprop_inner_product, src/vqcsim/parametric_circuit.cpp:180-279) over raw
// psi and lambda of one state: one work-group owns the same tile of both
// vectors, every dense / elementwise op of the pass (the inverted gates) is
// applied to both, and each OP_GRAD op adds Re sum conj(lambda) * (i P psi)
// over the tile, compensated, to partials[(gidx * nblk + tile) * 2 + {0, 1}].
// i P = R_P(pi) for the reference's R_P(theta) = exp(+i theta P / 2).
template <int TB>
__global__ __launch_bounds__(kThreads) void adjoint_pass_kernel(const PassArgs* __restrict__ pa, double2* __restrict__ psi,
                                                                double2* __restrict__ lam, double* __restrict__ partials,
                                                                uint64_t nblk, uint64_t tile0) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr uint32_t NT = 1u << TB;
  constexpr int EPT = (NT + kThreads - 1) / kThreads;
  __shared__ double red_s[kThreads / 64], red_c[kThreads / 64];
  double2* tp = reinterpret_cast<double2*>(smem);
  double2* tl = tp + NT;
  double2* ptab = tl + NT;
  const PassArgs& a = *pa;
  const TileMapDev tm{a.t, a.c, a.nE, a.E};
  const uint64_t tau = tile0 + blockIdx.x;
  const uint64_t base = tm.base(tau);
  uint64_t offs[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const uint32_t p = threadIdx.x + e * kThreads;
    offs[e] = base + tm.off(p & (NT - 1u));
    if (NT >= kThreads || p < NT) {
      tp[p] = psi[offs[e]];
      tl[p] = lam[offs[e]];
    }
  }
  __syncthreads();
  for (uint32_t o = a.op_begin; o < a.op_end; ++o) {
    const Op& op = a.ops[o];
    if (op.kind == OP_DENSE) {
      switch (op.k) {
        case 1: dense_op<1, TB>(op, tp, base, a.coefs); dense_op<1, TB>(op, tl, base, a.coefs); break;
        case 2: dense_op<2, TB>(op, tp, base, a.coefs); dense_op<2, TB>(op, tl, base, a.coefs); break;
        case 3: dense_op<3, TB>(op, tp, base, a.coefs); dense_op<3, TB>(op, tl, base, a.coefs); break;
        default: break;
      }
    } else if (op.kind == OP_ELEM) {
      elem_op<TB>(op, tp, base, a.terms, a.diags, a.coefs, ptab);
      __syncthreads();
      elem_op<TB>(op, tl, base, a.terms, a.diags, a.coefs, ptab);
    } else {
      const uint32_t fx = (uint32_t)op.gfx, zt = (uint32_t)op.gzt;
      const int kr = (int)((1u + op.nY) & 3u);  // i * i^nY = i^kr
      const bool bneg = __popcll(base & op.gzb) & 1;
      double s = 0.0, c = 0.0;
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const uint32_t p = threadIdx.x + e * kThreads;
        if (NT >= kThreads || p < NT) {
          const uint32_t q = p ^ fx;
          const double2 l = tl[p], x = tp[q];
          // conj(l) * x
          double re = l.x * x.x + l.y * x.y, im = l.x * x.y - l.y * x.x;
          const bool neg = bneg ^ (bool)(__popc(q & zt) & 1);
          re = neg ? -re : re;
          im = neg ? -im : im;
          const double v = kr == 0 ? re : kr == 1 ? -im : kr == 2 ? -re : im;
          dd_add(s, c, v);
        }
      }
      const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
      for (int d = 32; d >= 1; d >>= 1) {
        const double s2 = __shfl_xor(s, d, 64), c2 = __shfl_xor(c, d, 64);
        dd_merge(s, c, s2, c2);
      }
      if (lane == 0) { red_s[wave] = s; red_c[wave] = c; }
      __syncthreads();
      if (threadIdx.x == 0) {
        double ss = red_s[0], cc = red_c[0];
        for (int q = 1; q < kThreads / 64; ++q) dd_merge(ss, cc, red_s[q], red_c[q]);
        partials[((uint64_t)op.gidx * nblk + tau) * 2] = ss;
        partials[((uint64_t)op.gidx * nblk + tau) * 2 + 1] = cc;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const uint32_t p = threadIdx.x + e * kThreads;
    if (NT >= kThreads || p < NT) {
      psi[offs[e]] = tp[p];
      lam[offs[e]] = tl[p];
    }
  }
}

// out[point] = compensated sum over nblk tiles of the pass's partials
// (one work-group per gradient point)
// Compensated sum of the (s, c) pairs P[begin, end) by one block; thread 0
// stores the result pair at out[0..1].
__device__ __forceinline__ void block_pair_sum(const double* __restrict__ P, uint64_t begin, uint64_t end,
                                               double* __restrict__ out) {
  __shared__ double red_s[kThreads / 64], red_c[kThreads / 64];
  double s = 0.0, c = 0.0;
  for (uint64_t i = begin + threadIdx.x; i < end; i += kThreads) dd_merge(s, c, P[2 * i], P[2 * i + 1]);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int d = 32; d >= 1; d >>= 1) {
    const double s2 = __shfl_xor(s, d, 64), c2 = __shfl_xor(c, d, 64);
    dd_merge(s, c, s2, c2);
  }
  if (lane == 0) { red_s[wave] = s; red_c[wave] = c; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ss = red_s[0], cc = red_c[0];
    for (int q = 1; q < kThreads / 64; ++q) dd_merge(ss, cc, red_s[q], red_c[q]);
    out[0] = ss;
    out[1] = cc;
  }
}

// one block per point: out[2 p .. 2 p + 1] = sum of the nblk pairs of point p
__global__ __launch_bounds__(kThreads) void grad_reduce_kernel(const double* __restrict__ partials, uint64_t nblk,
                                                               double* __restrict__ out) {
  block_pair_sum(partials + (uint64_t)blockIdx.x * nblk * 2, 0, nblk, out + 2 * (uint64_t)blockIdx.x);
}

// first level for long partial lists: block (x, y) sums chunk x of point y
// into stage[2 (y * gridDim.x + x) ..]
__global__ __launch_bounds__(kThreads) void grad_reduce_chunks_kernel(const double* __restrict__ partials,
                                                                      uint64_t nblk, uint64_t chunk,
                                                                      double* __restrict__ stage) {
  const uint64_t b = (uint64_t)blockIdx.x * chunk, e = std::min<uint64_t>(nblk, b + chunk);
  block_pair_sum(partials + (uint64_t)blockIdx.y * nblk * 2, b, e,
                 stage + 2 * ((uint64_t)blockIdx.y * gridDim.x + blockIdx.x));
}

// ------------------------------------------------------ block-wise sampling
// Per-block probability mass for CompressedQuantumState::sampling_blockwise
// (src/cppsim/compressed_state.hpp:1612-1622): block (x, y) adds |a|^2 over
// chunk x of decoded block y (compensated) into stage pairs; a
// grad_reduce_kernel launch then folds the chunks of each block.
__global__ __launch_bounds__(kThreads) void block_sumsq_kernel(const double2* __restrict__ a, uint64_t namps,
                                                               uint64_t chunk, double* __restrict__ stage) {
  __shared__ double red_s[kThreads / 64], red_c[kThreads / 64];
  const double2* A = a + (uint64_t)blockIdx.y * namps;
  const uint64_t b = (uint64_t)blockIdx.x * chunk, e = std::min<uint64_t>(namps, b + chunk);
  double s = 0.0, c = 0.0;
  for (uint64_t i = b + threadIdx.x; i < e; i += kThreads) {
    const double2 v = A[i];
    dd_add(s, c, fma(v.x, v.x, v.y * v.y));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int d = 32; d >= 1; d >>= 1) {
    const double s2 = __shfl_xor(s, d, 64), c2 = __shfl_xor(c, d, 64);
    dd_merge(s, c, s2, c2);
  }
  if (lane == 0) { red_s[wave] = s; red_c[wave] = c; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ss = red_s[0], cc = red_c[0];
    for (int q = 1; q < kThreads / 64; ++q) dd_merge(ss, cc, red_s[q], red_c[q]);
    double* o = stage + 2 * ((uint64_t)blockIdx.y * gridDim.x + blockIdx.x);
    o[0] = ss;
    o[1] = cc;
  }
}

// Cumulative distribution of one decoded block in two levels
// (compressed_state.hpp:1654-1656 builds it sequentially): each workgroup
// scans kScanTile consecutive |a|^2 into cdf_local (inclusive, within the
// tile) and stores its tile total; cdf[j] = tile_prefix[tile(j)] + cdf_local[j].
constexpr uint32_t kScanPer = 16, kScanTile = kThreads * kScanPer;
__global__ __launch_bounds__(kThreads) void block_scan_kernel(const double2* __restrict__ a, uint64_t namps,
                                                              double* __restrict__ cdf_local,
                                                              double* __restrict__ tile_sum) {
  __shared__ double sh[kThreads];
  const uint64_t t0 = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanPer;
  double loc[kScanPer];
  double run = 0.0;
#pragma unroll
  for (uint32_t k = 0; k < kScanPer; ++k) {
    const uint64_t i = t0 + k;
    double p = 0.0;
    if (i < namps) {
      const double2 v = a[i];
      p = fma(v.x, v.x, v.y * v.y);
    }
    run += p;
    loc[k] = run;
  }
  // exclusive scan of the thread totals (Hillis-Steele in LDS)
  sh[threadIdx.x] = run;
  __syncthreads();
  for (uint32_t d = 1; d < kThreads; d <<= 1) {
    const double add = threadIdx.x >= d ? sh[threadIdx.x - d] : 0.0;
    __syncthreads();
    sh[threadIdx.x] += add;
    __syncthreads();
  }
  const double off = threadIdx.x ? sh[threadIdx.x - 1] : 0.0;
#pragma unroll
  for (uint32_t k = 0; k < kScanPer; ++k) {
    const uint64_t i = t0 + k;
    if (i < namps) cdf_local[i] = off + loc[k];
  }
  if (threadIdx.x == kThreads - 1) tile_sum[blockIdx.x] = sh[kThreads - 1];
}

// One thread: tile_incl[t] = sum of tile_sum[0..t] (in order); total at the end.
__global__ void block_scan_tiles_kernel(const double* __restrict__ tile_sum, uint32_t ntile,
                                        double* __restrict__ tile_incl) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double acc = 0.0;
  for (uint32_t t = 0; t < ntile; ++t) {
    acc += tile_sum[t];
    tile_incl[t] = acc;
  }
}

// One thread per shot: r = u * total, index = first j with cdf[j] >= r
// (std::lower_bound, compressed_state.hpp:1659-1662), clamped to the block.
__global__ void block_search_kernel(const double* __restrict__ cdf_local, const double* __restrict__ tile_incl,
                                    uint32_t ntile, uint64_t namps, const double* __restrict__ u, uint32_t nshots,
                                    uint64_t base, unsigned long long* __restrict__ out) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nshots) return;
  const double r = u[s] * tile_incl[ntile - 1];
  // first tile whose inclusive total reaches r
  uint32_t lo = 0, hi = ntile;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (tile_incl[mid] < r) lo = mid + 1;
    else hi = mid;
  }
  uint64_t j;
  if (lo >= ntile) {
    j = namps;
  } else {
    const double off = lo ? tile_incl[lo - 1] : 0.0;
    uint64_t a = (uint64_t)lo * kScanTile, b = std::min<uint64_t>(namps, a + kScanTile);
    while (a < b) {
      const uint64_t mid = (a + b) >> 1;
      if (off + cdf_local[mid] < r) a = mid + 1;
      else b = mid;
    }
    j = a;
  }
  if (j >= namps) j = namps - 1;
  out[s] = base + j;
}

// acc[map[k]] += pt[k] (compensated pairs) for the gradient points of one
// pass, in order: several points of a shared parameter add up
// deterministically.  One thread; a pass has at most a few hundred points.
__global__ void accumulate_points_kernel(const double* __restrict__ pt, uint32_t np, const int32_t* __restrict__ map,
                                         double* __restrict__ acc) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (uint32_t k = 0; k < np; ++k) {
    double* a = acc + 2 * (uint64_t)map[k];
    double s = a[0], c = a[1];
    dd_merge(s, c, pt[2 * k], pt[2 * k + 1]);
    a[0] = s;
    a[1] = c;
  }
}

// out[x] = sum_t c_t (P_t a)[x] for a Pauli sum in one sweep
// (Observable::apply_to_state, src/cppsim/pauli_operator.cpp:132-183): terms
// sharing a flip mask f share the load of a[x ^ f]; (P a)[x] = i^nY
// (-1)^{pc((x^f) & zy)} a[x^f] with i^nY folded into the coefficient.  Also
// accumulates Re sum conj(a[x]) out[x] into partials[blockIdx] (compensated).
// STORE = false: <a|H|a> only (Observable::get_expectation_value,
// src/cppsim/pauli_operator.cpp): nothing is stored, and the block writes
// compensated (re, im) partials as four doubles.
template <bool STORE>
__global__ __launch_bounds__(kThreads) void pauli_sum_kernel(const double2* __restrict__ a, double2* __restrict__ out,
                                                             uint64_t n, const PauliGroup* __restrict__ groups,
                                                             uint32_t ng, const PauliTermZ* __restrict__ terms,
                                                             double* __restrict__ partials) {
  __shared__ double red_s[kThreads / 64], red_c[kThreads / 64];
  __shared__ double red_s2[kThreads / 64], red_c2[kThreads / 64];
  double s = 0.0, c = 0.0, s2 = 0.0, c2 = 0.0;
  for (uint64_t x = (uint64_t)blockIdx.x * kThreads + threadIdx.x; x < n; x += (uint64_t)gridDim.x * kThreads) {
    double re = 0.0, im = 0.0;
    for (uint32_t g = 0; g < ng; ++g) {
      const PauliGroup G = groups[g];
      const uint64_t y = x ^ G.flip;
      // the group's coefficient at y: the terms' coefficients with the signs
      // of their Z bits.  Term fields are wave-uniform, so terms without Z
      // bits and real coefficients take branches that skip the parity and
      // the imaginary half.
      double fr = 0.0, fi = 0.0;
      bool cplx = false;
      for (uint32_t t = G.t0; t < G.t0 + G.nt; ++t) {
        const PauliTermZ T = terms[t];
        cplx |= T.ci != 0.0;
        if (T.z == 0ull) {
          fr += T.cr;
          fi += T.ci;
          continue;
        }
        const double sg = (__popcll(y & T.z) & 1) ? -1.0 : 1.0;
        fr = fma(sg, T.cr, fr);
        if (T.ci != 0.0) fi = fma(sg, T.ci, fi);
      }
      const double2 v = a[y];
      if (cplx) {
        re += fr * v.x - fi * v.y;
        im += fr * v.y + fi * v.x;
      } else {
        re = fma(fr, v.x, re);
        im = fma(fr, v.y, im);
      }
    }
    const double2 ax = a[x];
    if constexpr (STORE) {
      out[x] = make_double2(re, im);
    } else {
      dd_add(s2, c2, ax.x * im - ax.y * re);
    }
    dd_add(s, c, ax.x * re + ax.y * im);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int d = 32; d >= 1; d >>= 1) {
    double t1 = __shfl_xor(s, d, 64), t2 = __shfl_xor(c, d, 64);
    dd_merge(s, c, t1, t2);
    if constexpr (!STORE) {
      t1 = __shfl_xor(s2, d, 64);
      t2 = __shfl_xor(c2, d, 64);
      dd_merge(s2, c2, t1, t2);
    }
  }
  if (lane == 0) {
    red_s[wave] = s;
    red_c[wave] = c;
    red_s2[wave] = s2;
    red_c2[wave] = c2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ss = red_s[0], cc = red_c[0], ss2 = red_s2[0], cc2 = red_c2[0];
    for (int q = 1; q < kThreads / 64; ++q) {
      dd_merge(ss, cc, red_s[q], red_c[q]);
      dd_merge(ss2, cc2, red_s2[q], red_c2[q]);
    }
    if constexpr (STORE) {
      partials[2 * blockIdx.x] = ss;
      partials[2 * blockIdx.x + 1] = cc;
    } else {
      partials[4 * blockIdx.x] = ss;
      partials[4 * blockIdx.x + 1] = cc;
      partials[4 * blockIdx.x + 2] = ss2;
      partials[4 * blockIdx.x + 3] = cc2;
    }
  }
}

// ------------------------------------------------------- kernel launchers
template <int IN, int OUT, int TB, bool BIG = false>
static hipError_t launch_pass_tb(const PassArgs* d_args, const PassDyn& dy, uint64_t ntiles, hipStream_t st) {
  size_t lds = sizeof(double2) * ((1u << TB) + kPhaseTab);
  if (IN == IN_DEC || OUT == OUT_ENC) lds += sizeof(CodecScratch);
  static bool attr_set = false;
  if (!attr_set) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&pass_kernel<IN, OUT, TB, BIG>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr_set = true;
  }
  PassDyn d = dy;
  const uint64_t maxg = max_launch_groups();
  for (uint64_t t0 = 0; t0 < ntiles; t0 += maxg) {
    const uint64_t cnt = (ntiles - t0 < maxg) ? (ntiles - t0) : maxg;
    d.tile0 = dy.tile0 + t0;
    hipLaunchKernelGGL((pass_kernel<IN, OUT, TB, BIG>), dim3((unsigned)cnt), dim3(kThreads), lds, st, d_args, d);
  }
  return hipGetLastError();
}

template <int IN, int OUT, bool BIG>
static hipError_t launch_pass_mode(int tb, const PassArgs* d_args, const PassDyn& dy, uint64_t ntiles, hipStream_t st) {
  constexpr bool RR = IN == IN_RAW && OUT == OUT_RAW;
  switch (tb) {
    case 1: if constexpr (RR && !BIG) return launch_pass_tb<IN, OUT, 1, BIG>(d_args, dy, ntiles, st); break;
    case 2: if constexpr (RR && !BIG) return launch_pass_tb<IN, OUT, 2, BIG>(d_args, dy, ntiles, st); break;
    case 3: if constexpr (RR && !BIG) return launch_pass_tb<IN, OUT, 3, BIG>(d_args, dy, ntiles, st); break;
    case 4: if constexpr (RR) return launch_pass_tb<IN, OUT, 4, BIG>(d_args, dy, ntiles, st); break;
    case 5: return launch_pass_tb<IN, OUT, 5, BIG>(d_args, dy, ntiles, st);
    case 6: return launch_pass_tb<IN, OUT, 6, BIG>(d_args, dy, ntiles, st);
    case 7: return launch_pass_tb<IN, OUT, 7, BIG>(d_args, dy, ntiles, st);
    case 8: return launch_pass_tb<IN, OUT, 8, BIG>(d_args, dy, ntiles, st);
    case 9: return launch_pass_tb<IN, OUT, 9, BIG>(d_args, dy, ntiles, st);
    case 10: return launch_pass_tb<IN, OUT, 10, BIG>(d_args, dy, ntiles, st);
    case 11: return launch_pass_tb<IN, OUT, 11, BIG>(d_args, dy, ntiles, st);
    case 12: return launch_pass_tb<IN, OUT, 12, BIG>(d_args, dy, ntiles, st);
    case 13: return launch_pass_tb<IN, OUT, 13, BIG>(d_args, dy, ntiles, st);
    default: break;
  }
  return hipErrorInvalidValue;
}

template <bool BIG>
static hipError_t launch_pass_b(int in_mode, int out_mode, int tb, const PassArgs* d_args, const PassDyn& dy,
                                uint64_t ntiles, hipStream_t st) {
  if (in_mode == IN_RAW && out_mode == OUT_RAW) return launch_pass_mode<IN_RAW, OUT_RAW, BIG>(tb, d_args, dy, ntiles, st);
  if (in_mode == IN_DEC && out_mode == OUT_RAW) return launch_pass_mode<IN_DEC, OUT_RAW, BIG>(tb, d_args, dy, ntiles, st);
  if (in_mode == IN_RAW && out_mode == OUT_ENC) return launch_pass_mode<IN_RAW, OUT_ENC, BIG>(tb, d_args, dy, ntiles, st);
  return launch_pass_mode<IN_DEC, OUT_ENC, BIG>(tb, d_args, dy, ntiles, st);
}

// big selects the instantiation that also carries 4- and 5-target dense ops
hipError_t launch_pass(int in_mode, int out_mode, int tb, bool big, const PassArgs* d_args, const PassDyn& dy,
                       uint64_t ntiles, hipStream_t st) {
  return big ? launch_pass_b<true>(in_mode, out_mode, tb, d_args, dy, ntiles, st)
             : launch_pass_b<false>(in_mode, out_mode, tb, d_args, dy, ntiles, st);
}

static unsigned grid_for(uint64_t n, unsigned maxb = 65536u) {
  uint64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > maxb) b = maxb;
  return (unsigned)b;
}

hipError_t launch_fill_basis(double2* v, uint64_t n, uint64_t idx, hipStream_t st) {
  hipLaunchKernelGGL(fill_basis_kernel, dim3(grid_for(n)), dim3(256), 0, st, v, n, idx);
  return hipGetLastError();
}
hipError_t launch_fill_u8(uint8_t* p, uint64_t n, uint8_t val, hipStream_t st) {
  hipLaunchKernelGGL(fill_u8_kernel, dim3(grid_for(n)), dim3(256), 0, st, p, n, val);
  return hipGetLastError();
}
hipError_t launch_fill_u64(unsigned long long* p, uint64_t n, unsigned long long val, hipStream_t st) {
  hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_for(n)), dim3(256), 0, st, p, n, val);
  return hipGetLastError();
}
hipError_t launch_scale(double2* v, uint64_t n, double s, hipStream_t st) {
  hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(256), 0, st, v, n, s);
  return hipGetLastError();
}
hipError_t launch_commit_null(uint8_t* nf, uint8_t* nz, uint64_t nb, hipStream_t st) {
  hipLaunchKernelGGL(commit_null_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st, nf, nz, nb);
  return hipGetLastError();
}

hipError_t launch_block_bytes(const uint64_t* desc, const uint8_t* nullf, uint64_t nb, uint64_t cpb,
                              unsigned long long* out, hipStream_t st) {
  hipLaunchKernelGGL(block_bytes_kernel, dim3(grid_for(nb, 4096u)), dim3(256), 0, st, desc, nullf, nb, cpb, out);
  return hipGetLastError();
}

// out: 2 + 2 kCursorBlocks words (the sum and the maximum, then the partials)
hipError_t launch_cursor_stats(const unsigned long long* cur, uint32_t n, unsigned long long* out, hipStream_t st) {
  hipLaunchKernelGGL(cursor_stats_kernel, dim3(kCursorBlocks), dim3(256), 0, st, cur, n, out + 2);
  hipLaunchKernelGGL(cursor_stats_final_kernel, dim3(1), dim3(64), 0, st, out + 2, out);
  return hipGetLastError();
}

unsigned reduce_grid(uint64_t n) { return grid_for(n, 2048u); }

hipError_t launch_gather_desc(const uint64_t* desc, const uint32_t* bmap, uint64_t nblk, uint64_t cpb, uint64_t* out,
                              hipStream_t st) {
  if (nblk == 0) return hipSuccess;
  hipLaunchKernelGGL(gather_desc_kernel, dim3(grid_for(nblk * cpb, 65536u)), dim3(256), 0, st, desc, bmap, nblk, cpb,
                     out);
  return hipGetLastError();
}

hipError_t launch_sv_norm(const double2* a, uint64_t n, double* partials, unsigned grid, hipStream_t st) {
  hipLaunchKernelGGL((sv_reduce_kernel<0>), dim3(grid), dim3(kThreads), 0, st, a, a, n, partials);
  return hipGetLastError();
}
hipError_t launch_sv_inner(const double2* a, const double2* b, uint64_t n, double* partials, unsigned grid,
                           hipStream_t st) {
  hipLaunchKernelGGL((sv_reduce_kernel<1>), dim3(grid), dim3(kThreads), 0, st, a, b, n, partials);
  return hipGetLastError();
}
hipError_t launch_sv_pauli(const double2* a, uint64_t n, uint64_t f, uint64_t z, double* partials, unsigned grid,
                           hipStream_t st) {
  hipLaunchKernelGGL(sv_pauli_kernel, dim3(grid), dim3(kThreads), 0, st, a, n, f, z, partials);
  return hipGetLastError();
}
hipError_t launch_sv_apply_pauli(const double2* a, double2* out, uint64_t n, uint64_t f, uint64_t z, double cr,
                                 double ci, int acc, hipStream_t st) {
  hipLaunchKernelGGL(sv_apply_pauli_kernel, dim3(grid_for(n)), dim3(256), 0, st, a, out, n, f, z, cr, ci, acc);
  return hipGetLastError();
}


// adjoint pass over 2^L amplitudes in tiles of 2^tb (tb in 4..12)
template <int TB>
static hipError_t launch_adjoint_tb(const PassArgs* d_args, double2* psi, double2* lam, double* partials,
                                    uint64_t ntiles, hipStream_t st) {
  const size_t shm = sizeof(double2) * ((2ull << TB) + kPhaseTab);
  static bool attr = false;
  if (!attr && shm > 65536) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&adjoint_pass_kernel<TB>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return e;
  }
  attr = true;
  const uint64_t maxg = max_launch_groups();
  for (uint64_t t0 = 0; t0 < ntiles; t0 += maxg) {
    const unsigned cnt = (unsigned)std::min<uint64_t>(ntiles - t0, maxg);
    hipLaunchKernelGGL(adjoint_pass_kernel<TB>, dim3(cnt), dim3(kThreads), shm, st, d_args, psi, lam, partials, ntiles, t0);
  }
  return hipGetLastError();
}

hipError_t launch_adjoint_pass(int tb, const PassArgs* d_args, double2* psi, double2* lam, double* partials,
                               uint64_t ntiles, hipStream_t st) {
  switch (tb) {
    case 4: return launch_adjoint_tb<4>(d_args, psi, lam, partials, ntiles, st);
    case 5: return launch_adjoint_tb<5>(d_args, psi, lam, partials, ntiles, st);
    case 6: return launch_adjoint_tb<6>(d_args, psi, lam, partials, ntiles, st);
    case 7: return launch_adjoint_tb<7>(d_args, psi, lam, partials, ntiles, st);
    case 8: return launch_adjoint_tb<8>(d_args, psi, lam, partials, ntiles, st);
    case 9: return launch_adjoint_tb<9>(d_args, psi, lam, partials, ntiles, st);
    case 10: return launch_adjoint_tb<10>(d_args, psi, lam, partials, ntiles, st);
    case 11: return launch_adjoint_tb<11>(d_args, psi, lam, partials, ntiles, st);
    case 12: return launch_adjoint_tb<12>(d_args, psi, lam, partials, ntiles, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_accumulate_points(const double* pt, uint32_t np, const int32_t* map, double* acc, hipStream_t st) {
  if (np == 0) return hipSuccess;
  hipLaunchKernelGGL(accumulate_points_kernel, dim3(1), dim3(64), 0, st, pt, np, map, acc);
  return hipGetLastError();
}

// masses of nblocks consecutive decoded blocks of namps amplitudes into
// out[2 b .. 2 b + 1] (compensated pairs); stage holds nblocks *
// kGradReduceChunks pairs
hipError_t launch_block_sumsq(const double2* a, uint64_t namps, uint32_t nblocks, double* stage, double* out,
                              hipStream_t st) {
  if (nblocks == 0) return hipSuccess;
  const uint64_t nch = std::max<uint64_t>(1, std::min<uint64_t>(kGradReduceChunks, (namps + 4095) / 4096));
  const uint64_t chunk = (namps + nch - 1) / nch;
  hipLaunchKernelGGL(block_sumsq_kernel, dim3((unsigned)nch, nblocks), dim3(kThreads), 0, st, a, namps, chunk, stage);
  hipLaunchKernelGGL(grad_reduce_kernel, dim3(nblocks), dim3(kThreads), 0, st, (const double*)stage, nch, out);
  return hipGetLastError();
}

// tiles of the two-level block CDF
uint32_t block_scan_tiles(uint64_t namps) { return (uint32_t)((namps + kScanTile - 1) / kScanTile); }

// Samples of one decoded block: u holds nshots uniforms in [0, 1) on the
// device, out receives base + index per shot.  cdf_local: namps doubles;
// tile_sum, tile_incl: block_scan_tiles(namps) doubles each.
hipError_t launch_block_sample(const double2* a, uint64_t namps, double* cdf_local, double* tile_sum, double* tile_incl,
                               const double* u, uint32_t nshots, uint64_t base, unsigned long long* out,
                               hipStream_t st) {
  const uint32_t nt = block_scan_tiles(namps);
  hipLaunchKernelGGL(block_scan_kernel, dim3(nt), dim3(kThreads), 0, st, a, namps, cdf_local, tile_sum);
  hipLaunchKernelGGL(block_scan_tiles_kernel, dim3(1), dim3(64), 0, st, (const double*)tile_sum, nt, tile_incl);
  hipLaunchKernelGGL(block_search_kernel, dim3((nshots + 255) / 256), dim3(256), 0, st, (const double*)cdf_local,
                     (const double*)tile_incl, nt, namps, u, nshots, base, out);
  return hipGetLastError();
}

// stage: room for npoints * kGradReduceChunks pairs, or null for one level
hipError_t launch_grad_reduce(const double* partials, uint64_t nblk, uint32_t npoints, double* out, double* stage,
                              hipStream_t st) {
  if (npoints == 0) return hipSuccess;
  const uint64_t nch = std::min<uint64_t>(kGradReduceChunks, (nblk + 4095) / 4096);
  if (stage == nullptr || nch < 2) {
    hipLaunchKernelGGL(grad_reduce_kernel, dim3(npoints), dim3(kThreads), 0, st, partials, nblk, out);
    return hipGetLastError();
  }
  const uint64_t chunk = (nblk + nch - 1) / nch;
  hipLaunchKernelGGL(grad_reduce_chunks_kernel, dim3((unsigned)nch, npoints), dim3(kThreads), 0, st, partials, nblk,
                     chunk, stage);
  hipLaunchKernelGGL(grad_reduce_kernel, dim3(npoints), dim3(kThreads), 0, st, (const double*)stage, nch, out);
  return hipGetLastError();
}

hipError_t launch_pauli_sum(const double2* a, double2* out, uint64_t n, const PauliGroup* groups, uint32_t ng,
                            const PauliTermZ* terms, double* partials, unsigned grid, hipStream_t st) {
  hipLaunchKernelGGL(pauli_sum_kernel<true>, dim3(grid), dim3(kThreads), 0, st, a, out, n, groups, ng, terms, partials);
  return hipGetLastError();
}

// <a|H|a>: partials receive four doubles per block (re and im, compensated)
hipError_t launch_pauli_expect(const double2* a, uint64_t n, const PauliGroup* groups, uint32_t ng,
                               const PauliTermZ* terms, double* partials, unsigned grid, hipStream_t st) {
  hipLaunchKernelGGL(pauli_sum_kernel<false>, dim3(grid), dim3(kThreads), 0, st, a, (double2*)nullptr, n, groups, ng,
                     terms, partials);
  return hipGetLastError();
}

// ---------------------------------------------- allocation-free expectation
// One static device array per device, held until the process ends: the
// allocation-free Pauli expectation keeps its term tables, partials and
// final sums here instead of allocating.  In host code `scratch` names a
// host-side stand-in; kernels receive the device address from
// noalloc_scratch(), looked up once per device.
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// Device address of `scratch` on the current device (hipGetSymbolAddress on
// the first call per device, with that device current; kept per device).
hipError_t noalloc_scratch(void** out) {
  static std::mutex mu;
  static void* ptr[kNoAllocMaxDevices] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    ptr[dev] = p;
  }
  *out = ptr[dev];
  return hipSuccess;
}

// The block partials (s_re, c_re, s_im, c_im per block) merged in block
// order by one thread, with the operations of the host's compensated merge
// (two-sum of the running sum and the partial's sum, then its correction),
// so the result equals the host reduction of the same partials bit for bit.
__device__ __forceinline__ void dd_merge_seq(double& s, double& c, double s2, double c2) {
  const double t = s + s2;
  const double bp = t - s;
  c += (s - (t - bp)) + (s2 - bp);
  s = t;
  c += c2;
}
__global__ void merge_partials_kernel(const double* __restrict__ partials, unsigned grid, double* __restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double s0 = 0.0, c0 = 0.0, s1 = 0.0, c1 = 0.0;
  for (unsigned b = 0; b < grid; ++b) {
    dd_merge_seq(s0, c0, partials[4 * b], partials[4 * b + 1]);
    dd_merge_seq(s1, c1, partials[4 * b + 2], partials[4 * b + 3]);
  }
  out[0] = s0;
  out[1] = c0;
  out[2] = s1;
  out[3] = c1;
}

// <a|H|a> into scratch (its term tables already uploaded): block partials,
// then the four final doubles at kNoAllocFinalOff.
hipError_t launch_pauli_expect_noalloc(const double2* a, uint64_t n, uint32_t ng, unsigned char* scr, unsigned grid,
                                       hipStream_t st) {
  if (grid == 0 || grid > kNoAllocMaxGrid) return hipErrorInvalidValue;
  const PauliGroup* groups = reinterpret_cast<const PauliGroup*>(scr);
  const PauliTermZ* terms = reinterpret_cast<const PauliTermZ*>(scr + kNoAllocTermsOff);
  double* partials = reinterpret_cast<double*>(scr + kNoAllocPartialsOff);
  hipLaunchKernelGGL(pauli_sum_kernel<false>, dim3(grid), dim3(kThreads), 0, st, a, (double2*)nullptr, n, groups, ng,
                     terms, partials);
  hipLaunchKernelGGL(merge_partials_kernel, dim3(1), dim3(64), 0, st, partials, grid,
                     reinterpret_cast<double*>(scr + kNoAllocFinalOff));
  return hipGetLastError();
}

}  // namespace qsim
