// Shared layout between the C++ host planner and the HIP kernels.
//
// A "pass" streams a (group-local) state vector through LDS in tiles of 2^t
// amplitudes.  Tile element p (t bits) maps to local amplitude
//   base(tau) | off(p),  off(p) = (p & (2^c-1)) | sum_i bit(p, c+i) << E[i]
// where the low c local qubits are contiguous in every tile row and E[] are the
// extra local qubits that the pass needs in the tile.  Every op of the pass
// then works on LDS only.
//
// Compressed block store (per block of 2^B amplitudes = 2^(B+1) doubles):
//   chunk  = 64 consecutive doubles (32 amplitudes), one double per lane of a
//            wave64;
//   desc   = u64 per chunk: bits 0..39 payload offset in 8-byte words,
//            bits 40..47 width w (0 = all-zero chunk, 64 = verbatim doubles),
//            bits 48..63 signed step exponent e (value = k * 2^e);
//   payload= 64 lanes x w bits, lane l's zig-zag code at bit l*w (LSB first,
//            little-endian 64-bit words), 8*w bytes per chunk.
#pragma once
#include <stdint.h>

namespace qsim {

constexpr int kMaxDenseK = 5;      // non-diagonal targets of one dense op
constexpr int kMaxSel = 6;         // selector qubits (diagonal targets + controls)
constexpr int kMaxTileBits = 13;   // 2^13 complex128 = 128 KiB of LDS
constexpr int kThreads = 256;      // work-group size of the pass kernels
constexpr int kChunkAmpsLog2 = 5;  // 32 amplitudes = 64 doubles per codec chunk
constexpr int kMaxStoreSwaps = 16;     // SWAP gates folded into one ENC store
constexpr int kSubArenas = 64;         // allocation granularity of arena capacities
constexpr int kMaxSubArenas = 65536;   // allocation cursors per arena (the launch uses a power of two)
constexpr int kMaxLocalBits = 48;

enum OpKind : uint32_t { OP_DENSE = 0, OP_ELEM = 1, OP_GRAD = 2 };
enum InMode : int { IN_RAW = 0, IN_DEC = 1 };
enum OutMode : int { OUT_RAW = 0, OUT_ENC = 1 };
enum ZeroMode : int { ZERO_EXACT = 0, ZERO_THRESHOLD = 1 };
enum Engine : uint32_t { ENGINE_LDS = 0, ENGINE_WAVE = 1 };
constexpr uint32_t kWaveTile = 10;  // wave tile bits: 6 lane bits + 4 register slots
constexpr uint32_t kWaveTileSmall = 9;  // 6 lane bits + 3 register slots (8 rows per lane)
// A dispatch packet gives the grid in work-items as 32-bit values: launches of
// 256-thread work-groups are split at 2^23 groups (2^31 work-items).
// first-level chunks per gradient point in the two-level partial reduction
constexpr uint32_t kGradReduceChunks = 64;
constexpr uint64_t kMaxLaunchBlocks = 1ull << 23;

constexpr uint64_t kNoPhaseTable = ~0ull;

// One op of a pass.  Dense: 2^nsel matrices of 2^k x 2^k (row-major, matrix
// index bit i <-> tile bit tpos[i]); the matrix is selected by the selector
// bits (each a tile bit or a base bit).  Elementwise: a run of diagonal gates,
// given as phase terms (angle added when the condition holds) and diagonal
// tables.
struct Op {
  uint32_t kind;
  uint32_t k;
  uint32_t nsel;
  uint32_t mat_off;        // complex offset into the pass coefficient array
  uint64_t ident_mask;     // selector value s is the identity when bit s set
  uint16_t tpos[8];        // tile bit of target i
  uint16_t tsorted[8];     // ascending copy of tpos (zero insertion)
  uint16_t sel_pos[8];
  uint8_t sel_tile[8];     // 1: tile bit, 0: base (local index) bit
  uint32_t nterms, term_off;
  uint32_t ndiag, diag_off;
  // OP_GRAD (adjoint reverse sweep): partial slot gidx of the pass receives
  // Re sum conj(lambda) * (i P psi) over the tile, P the generator's Pauli
  // string: X/Y flips on tile bits gfx, Z/Y signs on tile bits gzt and base
  // bits gzb, nY Y factors
  uint32_t gidx, nY;
  uint64_t gfx, gzt, gzb;
  // OP_DENSE matrix structure (kDenseReal / kDenseRealDiagImagOff, else 0);
  // the one-target paths then skip the products with zero parts
  uint32_t real, pad_;
  // OP_ELEM on the wave engine: the op's phase terms, when ptab_off is not
  // kNoPhaseTable, are applied from factor tables at that complex offset of
  // the pass's phase-table array (the terms stay for the other engines)
  uint64_t ptab_off = kNoPhaseTable;
  uint32_t ptab_q;         // offsets from ptab_off (complex) of Q (U follows Q), G and M
  uint32_t ptab_g, ptab_m, pad2_;
  uint32_t ptab_slots;     // bit j: slot j has a factor; bit 8 + k: lane bit k has a G factor
  uint32_t ptab_flags;     // kPtabQ | kPtabU | kPtabM | kPtabRows | kPtabFt | kPtabK: those tables hold a term
  uint32_t ptab_nk;        // kPtabK: M's row is the value of these base bits (local qubits ptab_kpos[])
  uint8_t ptab_kpos[8];
};

// Phase tables of an elementwise op on a wave tile: lanes on tile bits 0..5,
// RB register slots on tile bits 6 .., tau the tile's index in its group
// (ntau = 2^(L - t) tiles).  Element (lane l, row v) of tile tau gets
//   Q[l] U[tau] M[tau][l] * prod_{j : bit j of v} (C[j][l] F[tau][j]) * R[v],
// each entry cis(sum of the angles of its terms, in term order).  A term on
// one slot j that requires s_j = 0 contributes +angle to the slot-free part
// and -angle to slot j's factor (angle [s_j = 0] = angle - angle s_j).
//   C[4][64]    single-slot terms without a base condition,
//   R[16]       terms on two or more slots (no lane or base condition),
//   F[ntau][4]  single-slot terms with a base condition (no lane condition),
//   Q[64]       slot-free parts without a base condition,
//   U[ntau]     slot-free parts with a base condition and no lane condition,
//   G[ntau][8]  (kPtabG, groups of more than 2^16 tiles) slot-free parts
//               with a base condition and a condition on one lane bit k,
//               kept as angles (real part): lane l gets cis of
//               the sum of G[tau][k] over its set bits k (a term requiring
//               bit k = 0 adds +angle to U, -angle to G),
//   M[ntau][64] the other slot-free parts with lane and base conditions
//               (only when ntau <= 2^16; M then takes every slot-free part
//               and Q, U are unused),
//   K[2^nk][64] (kPtabK) replaces M when the terms with lane and base
//               conditions test at most kPtabMaxKeyBits distinct base bits:
//               row = the value of those bits in the tile base (Q and U stay)
// at complex offsets 0, 256, 272 (F, 4 ntau entries when kPtabFt), ptab_q
// (Q, 64), ptab_q + 64 (U, ntau entries when kPtabU), ptab_g (G, 8 ntau when
// kPtabG) and ptab_m (M or K).  Ops whose terms do not separate so (a slot
// term with both a lane and a base condition) keep the per-term loop.
constexpr uint32_t kPtabC = 0, kPtabR = 256, kPtabF = 272;
constexpr uint32_t kPtabQ = 1, kPtabRows = 2, kPtabU = 4, kPtabM = 8;
// G replaces M for single-lane-bit terms when M would exceed its limit
constexpr uint32_t kPtabG = 16;
constexpr uint32_t kPtabFt = 32;  // the per-tile slot factors F hold a term
constexpr uint32_t kPtabK = 64;   // the lane table is keyed by base bits
constexpr uint32_t kPtabMaxKeyBits = 6;
constexpr uint32_t kPhaseTableMinTerms = 1;     // fewer terms: the per-term loop
constexpr uint32_t kPhaseTableMaxTauLog = 20;   // F and U of at most 2^20 tiles (80 MiB)
constexpr uint32_t kPhaseTableMaxTauLogM = 16;  // M of at most 2^16 x 64 entries (64 MiB)
constexpr uint64_t kPhaseTableMaxEntries = 1ull << 26;  // 1 GiB of tables per program

// Dense matrix structures with cheaper one-target paths: every entry real
// (RY, H, X), or one target with real diagonal and imaginary off-diagonal
// entries (RX)
constexpr uint32_t kDenseReal = 1;
constexpr uint32_t kDenseRealDiagImagOff = 2;

// Pauli-sum application: terms grouped by flip mask; coefficient of a term
// already includes i^nY and any overall scale
struct PauliGroup {
  uint64_t flip;
  uint32_t t0, nt;
};
struct PauliTermZ {
  uint64_t z;
  double cr, ci;
};

// Device scratch of the allocation-free Pauli expectation: the grouped term
// tables, four doubles of block partials per block of the reduction and the
// four final doubles, at fixed offsets in one static device array.
constexpr uint32_t kNoAllocMaxTerms = 1024;
constexpr uint32_t kNoAllocMaxGrid = 2048;  // the reductions' largest grid
constexpr uint64_t kNoAllocTermsOff = (uint64_t)kNoAllocMaxTerms * sizeof(PauliGroup);
constexpr uint64_t kNoAllocPartialsOff = kNoAllocTermsOff + (uint64_t)kNoAllocMaxTerms * sizeof(PauliTermZ);
constexpr uint64_t kNoAllocFinalOff = kNoAllocPartialsOff + (uint64_t)kNoAllocMaxGrid * 4 * sizeof(double);
constexpr uint64_t kNoAllocBytes = kNoAllocFinalOff + 4 * sizeof(double);
constexpr int kNoAllocMaxDevices = 64;

// 32 bytes: kernels read a term with one scalar load.  An elementwise op's
// term list is padded to a multiple of kTermGroup with terms that never
// match (bmask 0, bval 1), so terms are read kTermGroup at a time.
struct alignas(32) PhaseTerm {
  uint32_t tmask, tval;    // condition on the tile element index p
  uint64_t bmask, bval;    // condition on the tile base (local index)
  double angle;
};
constexpr uint32_t kTermGroup = 2;

struct DiagTable {
  uint32_t m;              // number of qubits
  uint32_t coef_off;       // complex offset, 2^m entries
  uint16_t pos[8];
  uint8_t tile[8];
};

struct PassArgs {
  // tile map
  uint32_t t, c, nE, L;
  uint32_t E[32];
  uint32_t op_begin, op_end;
  uint32_t maxk;           // largest dense op of the pass (4 or 5 select the BIG kernel)
  uint32_t engine;         // ENGINE_LDS: LDS tile kernel; ENGINE_WAVE: wave-resident kernel
  const Op* ops;
  const PhaseTerm* terms;
  const DiagTable* diags;
  const double* coefs;     // interleaved complex
  const double* ptab;      // phase tables of the program's elementwise ops (interleaved complex)
  // compressed side
  uint32_t B;              // block bits
  uint32_t khi;            // L - B: group-local high bits (blocks per group = 2^khi)
  uint64_t cpb;            // chunks per block
  double eps_rel, eps_abs;
  int guard;
  int zero_mode;
  double zero_thr;
  int e_abs;               // ilogb(2*eps_abs), INT_MIN when eps_abs == 0
  uint32_t ovf_hi;         // high word of the smallest |x| whose 2*fma(eps_rel,|x|,eps_abs)
                           // overflows (0x7ff00000 when none is finite)
  double p_abs;            // 2^e_abs
  // chunk permutation of an ENC store: the chunk at local position x is
  // written to position s_(n-1)(...s_0(x)), s_i swapping local bits
  // swp[2i] and swp[2i + 1] (all >= kChunkAmpsLog2): trailing SWAP gates
  uint32_t nswap;
  uint8_t swp[2 * kMaxStoreSwaps];
  // tile permutation of an ENC store on the wave engine (trailing SWAP gates
  // with a qubit below the chunk bits, both qubits tile bits): the value at
  // tile position 64 v + l is placed at the position the swaps send it to,
  // through the encoder's layout transpose, whose LDS slot for it is
  // tperm_lane[l] + tperm_row[v]
  uint32_t tperm;
  uint16_t tperm_row[16];
  uint16_t tperm_lane[64];
};

// Per-launch buffers of a pass (passed by value).
struct PassDyn {
  double* raw;              // group slot g at raw + 2*(g << L) doubles
  const uint32_t* bmap;     // [slot][2^khi] -> block id
  const uint8_t* null_in;   // [nb] null flags at stage start
  const uint64_t* desc_in;
  const uint64_t* arena_in;
  uint64_t* desc_out;
  uint64_t* arena_out;
  const uint64_t* sub_base; // sub-arena s spans words [sub_base[s], sub_base[s + 1])
  const uint32_t* slot_region;  // region store: group slot -> output region r (sub-arenas r (sub_mask + 1) + i,
                                //   bounds at sub_base[2s], sub_base[2s + 1]); else null
  unsigned long long* cursors;  // [sub_mask + 1] (stride 16 words)
  uint32_t sub_mask;        // number of sub-arenas - 1 (power of two)
  uint32_t wave_subs;       // persistent DEC->ENC: wave w of the grid owns sub-arena w
                            //   (as many sub-arenas as waves; cursors[16 w] = its demand)
  uint32_t desc_staged;     // desc_in holds the launch's blocks in bmap order: the block at
                            //   bmap position i (from this launch's bmap) at desc_in[i cpb ..]
  uint32_t pad_;
  uint32_t* overflow;
  uint8_t* nonzero;         // [nb]: set when a block null at start holds data
  uint64_t tile0;           // first tile of this launch
  // RAW->RAW wave passes: output rows go to raw_out when set (out of place)
  double* raw_out;
  // lambda-only reverse passes (raw holds lambda): the gradient points at the
  // start of the pass read psi at the same positions from `psi` (the stored
  // forward state) and write their tile sums to
  // partials[(gidx * part_tiles + tile) * 2]
  const double* psi;
  double* partials;
  uint64_t part_tiles;
};

// One instance of the persistent batch kernel: a small circuit run start to
// end by one work-group (forward passes on 1024-amplitude wave tiles,
// lambda = H psi with <psi|lambda>, then the reverse passes on
// 512-amplitude single-wave tiles with their gradient points), its work-group
// taking the next instance from a counter when done.
struct BatchJob {
  uint32_t n, nfwd, nrev, ng;
  const PassArgs* fwd;
  const PassArgs* rev;
  const PauliGroup* groups;
  const PauliTermZ* terms;
  const uint32_t* map_off;  // per reverse pass: first of its gradient points in pmap
  const uint32_t* map_cnt;  // per reverse pass: number of gradient points
  const int32_t* pmap;      // gradient point -> parameter
  double2* psi;
  double2* lam;
  double* acc;              // [0, 1] <psi|lambda> (compensated), [2 + 2p, 3 + 2p] gradient sums
};
constexpr uint32_t kBatchMaxPoints = 256;  // gradient points of one reverse pass
constexpr uint32_t kBatchMaxQubits = 16;   // largest instance of the persistent kernel
constexpr uint32_t kBatchMinQubits = 10;   // one 1024-amplitude forward tile

__host__ __device__ inline uint64_t desc_pack(uint64_t off_words, uint32_t w, int e) {
  return (off_words & ((1ull << 40) - 1)) | ((uint64_t)(w & 0xff) << 40) |
         ((uint64_t)(uint16_t)(int16_t)e << 48);
}
__host__ __device__ inline uint64_t desc_off(uint64_t d) { return d & ((1ull << 40) - 1); }
__host__ __device__ inline uint32_t desc_w(uint64_t d) { return (uint32_t)((d >> 40) & 0xff); }
__host__ __device__ inline int desc_e(uint64_t d) { return (int)(int16_t)(uint16_t)(d >> 48); }

}  // namespace qsim
