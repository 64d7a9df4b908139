// Host side of libqsim_hip.so: gate handling, circuit optimizer, LDS pass
// planner, compressed block store with its stage loop, and the C ABI
// declared in include/qsim.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <random>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "qsim.h"
#include "qsim_internal.h"

namespace qsim {


hipError_t launch_pass(int in_mode, int out_mode, int tb, bool big, const PassArgs* d_args, const PassDyn& dy,
                       uint64_t ntiles, hipStream_t st);
hipError_t launch_wave(int tile_bits, int in_mode, int out_mode, const PassArgs* d_args, const PassDyn& dy, uint64_t ntiles,
                       hipStream_t st);
hipError_t launch_fill_basis(double2* v, uint64_t n, uint64_t idx, hipStream_t st);
hipError_t launch_fill_u8(uint8_t* p, uint64_t n, uint8_t val, hipStream_t st);
hipError_t launch_fill_u64(unsigned long long* p, uint64_t n, unsigned long long val, hipStream_t st);
hipError_t launch_scale(double2* v, uint64_t n, double s, hipStream_t st);
hipError_t launch_commit_null(uint8_t* nf, uint8_t* nz, uint64_t nb, hipStream_t st);
unsigned reduce_grid(uint64_t n);
hipError_t noalloc_scratch(void** out);
hipError_t launch_pauli_expect_noalloc(const double2* a, uint64_t n, uint32_t ng, unsigned char* scr, unsigned grid,
                                       hipStream_t st);
hipError_t launch_sv_norm(const double2* a, uint64_t n, double* partials, unsigned grid, hipStream_t st);
hipError_t launch_sv_inner(const double2* a, const double2* b, uint64_t n, double* partials, unsigned grid,
                           hipStream_t st);
hipError_t launch_sv_pauli(const double2* a, uint64_t n, uint64_t f, uint64_t z, double* partials, unsigned grid,
                           hipStream_t st);
hipError_t launch_sv_apply_pauli(const double2* a, double2* out, uint64_t n, uint64_t f, uint64_t z, double cr,
                                 double ci, int acc, hipStream_t st);
// out: 2 + 128 words (sum and maximum, then per-block partials)
hipError_t launch_cursor_stats(const unsigned long long* cur, uint32_t n, unsigned long long* out, hipStream_t st);
hipError_t launch_adjoint_pass(int tb, const PassArgs* d_args, double2* psi, double2* lam, double* partials,
                               uint64_t ntiles, hipStream_t st);
hipError_t launch_grad_reduce(const double* partials, uint64_t nblk, uint32_t npoints, double* out, double* stage,
                              hipStream_t st);
hipError_t launch_accumulate_points(const double* pt, uint32_t np, const int32_t* map, double* acc, hipStream_t st);
hipError_t launch_block_sumsq(const double2* a, uint64_t namps, uint32_t nblocks, double* stage, double* out,
                              hipStream_t st);
uint32_t block_scan_tiles(uint64_t namps);
hipError_t launch_block_sample(const double2* a, uint64_t namps, double* cdf_local, double* tile_sum, double* tile_incl,
                               const double* u, uint32_t nshots, uint64_t base, unsigned long long* out,
                               hipStream_t st);
hipError_t launch_gather_desc(const uint64_t* desc, const uint32_t* bmap, uint64_t nblk, uint64_t cpb, uint64_t* out,
                              hipStream_t st);
hipError_t launch_batch_persistent(const BatchJob* jobs, uint32_t njobs, unsigned* next, unsigned groups,
                                   hipStream_t st);
// waves of the persistent DEC->ENC grid over ntiles tiles
uint64_t wave_persistent_waves(uint64_t ntiles);
hipError_t launch_phase_table(const PassArgs* d_args, const Op* d_op, uint64_t ntau, uint64_t entries, double* ptab,
                              hipStream_t st);
hipError_t launch_wave_adjoint(int tile_bits, const PassArgs* d_args, double2* psi, double2* lam, double* partials,
                               uint64_t ntiles, hipStream_t st);
hipError_t launch_wave_pair_adjoint(int tile_bits, const PassArgs* d_args, double2* psi, double2* lam,
                                    double* partials, uint64_t ntiles, hipStream_t st);
hipError_t launch_wave_lam(const PassArgs* d_args, const PassDyn& dy, uint64_t ntiles, hipStream_t st);
hipError_t launch_pauli_sum(const double2* a, double2* out, uint64_t n, const PauliGroup* groups, uint32_t ng,
                            const PauliTermZ* terms, double* partials, unsigned grid, hipStream_t st);
hipError_t launch_pauli_expect(const double2* a, uint64_t n, const PauliGroup* groups, uint32_t ng,
                               const PauliTermZ* terms, double* partials, unsigned grid, hipStream_t st);
hipError_t launch_block_bytes(const uint64_t* desc, const uint8_t* nullf, uint64_t nb, uint64_t cpb,
                              unsigned long long* out, hipStream_t st);

using cplx = std::complex<double>;
static thread_local std::string g_err;

// A failing runtime call also leaves its code as the thread's "last error",
// which the next launch check would report again; it is cleared before the
// throw so one failure is reported once.
#define HIPCHECK(x)                                                                                   \
  do {                                                                                                \
    hipError_t e_ = (x);                                                                              \
    if (e_ != hipSuccess) {                                                                           \
      (void)hipGetLastError();                                                                        \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " in " + #x);      \
    }                                                                                                 \
  } while (0)

// ------------------------------------------------------------------ context
static hipStream_t stream() {
  static hipStream_t st = [] {
    hipStream_t s;
    HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
  }();
  return st;
}

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t b) {
    release();
    if (b == 0) b = 16;
    HIPCHECK(hipMalloc(&p, b));
    bytes = b;
  }
  void ensure(size_t b) {
    if (b > bytes) alloc(b);
  }
  template <class T>
  T* as() const { return reinterpret_cast<T*>(p); }
};

static void upload(DBuf& d, const void* h, size_t b) {
  d.ensure(b);
  if (b) HIPCHECK(hipMemcpyAsync(d.p, h, b, hipMemcpyHostToDevice, stream()));
}

// -------------------------------------------------------------------- gates
struct HGate {
  std::vector<uint32_t> t, tprop, c, cval, pauli;
  std::vector<cplx> m;  // 2^k x 2^k row-major
  uint32_t flags = 0;
  int32_t param = -1;
  int64_t origin = -1;  // input index for gates the optimizer left unmerged
  bool parametric() const { return flags & QSIM_GATE_PARAMETRIC; }
};

static std::vector<HGate> parse_gates(const qsim_gate* g, uint64_t n, const double* pool, uint64_t pool_len) {
  std::vector<HGate> out(n);
  for (uint64_t i = 0; i < n; ++i) {
    const qsim_gate& q = g[i];
    if (q.n_targets > QSIM_MAX_GATE_QUBITS || q.n_controls > QSIM_MAX_GATE_QUBITS)
      throw std::invalid_argument("gate " + std::to_string(i) + ": too many qubits");
    HGate& h = out[i];
    h.t.assign(q.targets, q.targets + q.n_targets);
    h.tprop.assign(q.target_commute, q.target_commute + q.n_targets);
    h.pauli.assign(q.pauli_ids, q.pauli_ids + q.n_targets);
    h.c.assign(q.controls, q.controls + q.n_controls);
    h.cval.assign(q.control_values, q.control_values + q.n_controls);
    h.flags = q.flags;
    h.param = q.param_index;
    h.origin = (int64_t)i;
    const uint64_t dim = 1ull << q.n_targets;
    if (q.matrix_offset + dim * dim > pool_len / 2)
      throw std::invalid_argument("gate " + std::to_string(i) + ": matrix outside the pool");
    h.m.resize(dim * dim);
    for (uint64_t k = 0; k < dim * dim; ++k)
      h.m[k] = cplx(pool[2 * (q.matrix_offset + k)], pool[2 * (q.matrix_offset + k) + 1]);
    std::vector<uint32_t> all = h.t;
    all.insert(all.end(), h.c.begin(), h.c.end());
    std::sort(all.begin(), all.end());
    if (std::adjacent_find(all.begin(), all.end()) != all.end())
      throw std::invalid_argument("gate " + std::to_string(i) + ": repeated qubit");
  }
  return out;
}

static void check_qubits(const std::vector<HGate>& gs, uint32_t n) {
  for (size_t i = 0; i < gs.size(); ++i) {
    for (auto q : gs[i].t)
      if (q >= n) throw std::invalid_argument("gate " + std::to_string(i) + ": target out of range");
    for (auto q : gs[i].c)
      if (q >= n) throw std::invalid_argument("gate " + std::to_string(i) + ": control out of range");
  }
}

// ----------------------------------------------------------- the optimizer
// Commutation of two gates by their qubit roles (src/cppsim/gate.cpp:12-31,
// src/cppsim/qubit_info.cpp:4-31).
static bool gates_commute(const HGate& a, const HGate& b) {
  for (size_t i = 0; i < a.t.size(); ++i) {
    for (size_t j = 0; j < b.t.size(); ++j)
      if (a.t[i] == b.t[j] && (a.tprop[i] & b.tprop[j]) == 0) return false;
    for (size_t j = 0; j < b.c.size(); ++j)
      if (a.t[i] == b.c[j] && !(a.tprop[i] & QSIM_COMMUTE_Z)) return false;
  }
  for (size_t i = 0; i < a.c.size(); ++i)
    for (size_t j = 0; j < b.t.size(); ++j)
      if (a.c[i] == b.t[j] && !(b.tprop[j] & QSIM_COMMUTE_Z)) return false;
  return true;
}

// Matrix of gate g on the ordered target list nt (index bit i <-> nt[i]);
// g's controls that appear in nt act as conditions, the others are assumed
// satisfied.
static std::vector<cplx> extend_matrix(const HGate& g, const std::vector<uint32_t>& nt) {
  const size_t K = nt.size(), D = 1ull << K;
  auto pos = [&](uint32_t q) -> int {
    for (size_t i = 0; i < K; ++i)
      if (nt[i] == q) return (int)i;
    return -1;
  };
  std::vector<int> tpos(g.t.size());
  uint64_t gmask = 0;
  for (size_t i = 0; i < g.t.size(); ++i) {
    tpos[i] = pos(g.t[i]);
    gmask |= 1ull << tpos[i];
  }
  uint64_t cmask = 0, cval = 0;
  for (size_t i = 0; i < g.c.size(); ++i) {
    int p = pos(g.c[i]);
    if (p >= 0) {
      cmask |= 1ull << p;
      if (g.cval[i]) cval |= 1ull << p;
    }
  }
  const size_t gd = 1ull << g.t.size();
  std::vector<cplx> M(D * D, cplx(0, 0));
  for (uint64_t col = 0; col < D; ++col) {
    if ((col & cmask) != cval) {
      M[col * D + col] = 1.0;
      continue;
    }
    uint64_t cg = 0;
    for (size_t i = 0; i < g.t.size(); ++i) cg |= ((col >> tpos[i]) & 1ull) << i;
    for (uint64_t rg = 0; rg < gd; ++rg) {
      uint64_t row = col & ~gmask;
      for (size_t i = 0; i < g.t.size(); ++i) row |= ((rg >> i) & 1ull) << tpos[i];
      M[row * D + col] = g.m[rg * gd + cg];
    }
  }
  return M;
}

// gate::merge (src/cppsim/gate_merge.cpp:19-149, 316-356): qubit roles of the
// product, targets sorted, matrix = second * first.
static HGate merge_gates(const HGate& g1, const HGate& g2) {
  struct TI { uint32_t q, prop; };
  std::vector<TI> nt;
  std::vector<std::pair<uint32_t, uint32_t>> nc;
  auto find_t = [](const HGate& g, uint32_t q) -> int {
    for (size_t i = 0; i < g.t.size(); ++i)
      if (g.t[i] == q) return (int)i;
    return -1;
  };
  auto find_c = [](const HGate& g, uint32_t q) -> int {
    for (size_t i = 0; i < g.c.size(); ++i)
      if (g.c[i] == q) return (int)i;
    return -1;
  };
  for (size_t i = 0; i < g1.t.size(); ++i) {
    int j = find_t(g2, g1.t[i]);
    if (j >= 0) { nt.push_back({g1.t[i], g1.tprop[i] & g2.tprop[j]}); continue; }
    j = find_c(g2, g1.t[i]);
    if (j >= 0) { nt.push_back({g1.t[i], g1.tprop[i] & QSIM_COMMUTE_Z}); continue; }
    nt.push_back({g1.t[i], g1.tprop[i]});
  }
  for (size_t i = 0; i < g1.c.size(); ++i) {
    int j = find_t(g2, g1.c[i]);
    if (j >= 0) { nt.push_back({g1.c[i], g2.tprop[j] & QSIM_COMMUTE_Z}); continue; }
    j = find_c(g2, g1.c[i]);
    if (j >= 0) {
      if (g1.cval[i] == g2.cval[j]) nc.push_back({g1.c[i], g1.cval[i]});
      else nt.push_back({g1.c[i], QSIM_COMMUTE_Z});
      continue;
    }
    nt.push_back({g1.c[i], QSIM_COMMUTE_Z});
  }
  for (size_t i = 0; i < g2.t.size(); ++i) {
    if (find_t(g1, g2.t[i]) >= 0 || find_c(g1, g2.t[i]) >= 0) continue;
    nt.push_back({g2.t[i], g2.tprop[i]});
  }
  for (size_t i = 0; i < g2.c.size(); ++i) {
    if (find_t(g1, g2.c[i]) >= 0 || find_c(g1, g2.c[i]) >= 0) continue;
    nt.push_back({g2.c[i], QSIM_COMMUTE_Z});
  }
  std::sort(nt.begin(), nt.end(), [](const TI& a, const TI& b) { return a.q < b.q; });
  std::sort(nc.begin(), nc.end());
  HGate out;
  for (auto& x : nt) { out.t.push_back(x.q); out.tprop.push_back(x.prop); out.pauli.push_back(0); }
  for (auto& x : nc) { out.c.push_back(x.first); out.cval.push_back(x.second); }
  const auto M1 = extend_matrix(g1, out.t);
  const auto M2 = extend_matrix(g2, out.t);
  const size_t D = 1ull << out.t.size();
  out.m.assign(D * D, cplx(0, 0));
  for (size_t r = 0; r < D; ++r)
    for (size_t k = 0; k < D; ++k) {
      const cplx a = M2[r * D + k];
      if (a == cplx(0, 0)) continue;
      for (size_t c = 0; c < D; ++c) out.m[r * D + c] += a * M1[k * D + c];
    }
  out.flags = g1.flags & g2.flags;
  out.param = -1;
  out.origin = -1;
  return out;
}

// QuantumCircuitOptimizer::optimize(circuit, max_block_size, swap_level=0)
// (src/cppsim/circuit_optimizer.cpp:113-168, helpers 35-111).
static std::vector<HGate> optimize_circuit(std::vector<HGate> gl, uint32_t max_block) {
  auto rightmost = [&](size_t i) {
    size_t cur = i + 1;
    for (; cur < gl.size(); ++cur)
      if (!gates_commute(gl[i], gl[cur])) break;
    return cur - 1;
  };
  auto leftmost = [&](size_t i) {
    long cur = (long)i - 1;
    for (; cur >= 0; --cur)
      if (!gates_commute(gl[i], gl[cur])) break;
    return (size_t)(cur + 1);
  };
  auto merged_size = [&](size_t a, size_t b) {
    std::set<uint32_t> s;
    for (auto q : gl[a].t) s.insert(q);
    for (auto q : gl[b].t) s.insert(q);
    for (auto q : gl[a].c) s.insert(q);
    for (auto q : gl[b].c) s.insert(q);
    return (uint32_t)s.size();
  };
  bool merged = true;
  while (merged) {
    merged = false;
    for (size_t i1 = 0; i1 < gl.size(); ++i1) {
      for (size_t i2 = i1 + 1; i2 < gl.size(); ++i2) {
        if (gl[i1].parametric() || gl[i2].parametric()) continue;
        if (merged_size(i1, i2) > max_block) continue;
        if (leftmost(i2) > rightmost(i1) + 1) continue;
        HGate mg = merge_gates(gl[i1], gl[i2]);
        const size_t ins = std::max(leftmost(i2), i1 + 1);
        gl.erase(gl.begin() + i2);
        gl.insert(gl.begin() + ins, std::move(mg));
        gl.erase(gl.begin() + i1);
        --i2;
        merged = true;
      }
    }
  }
  return gl;
}

// ---------------------------------------------------------------- planner
// A gate rewritten in the local qubit numbering of one pass program.
struct LGate {
  // gradient point of the adjoint reverse sweep (parameter index), or -1; its
  // Pauli string has X/Y on gx and Z/Y on gz (qubit masks), nY Y factors
  int32_t grad = -1;
  uint64_t gx = 0, gz = 0;
  uint32_t nY = 0;
  bool diag = false;
  // a trailing SWAP folded into the encoder's tile permutation: it holds its
  // qubits in the tile while the pass is planned and emits no op
  bool noop = false;
  // index of the circuit gate this op inverts in a reverse program, else -1
  int64_t orig = -1;
  std::vector<uint32_t> N;      // non-diagonal targets (matrix bit order)
  std::vector<uint32_t> S;      // selector qubits
  std::vector<cplx> mats;       // 2^|S| matrices of 2^|N| x 2^|N|
  uint64_t ident = 0;
  std::vector<uint32_t> Q;      // diagonal form
  std::vector<cplx> d;
};

static bool is_identity(const cplx* m, size_t D) {
  for (size_t r = 0; r < D; ++r)
    for (size_t c = 0; c < D; ++c)
      if (m[r * D + c] != (r == c ? cplx(1, 0) : cplx(0, 0))) return false;
  return true;
}

static LGate lower_gate(const HGate& g, const std::vector<int>& qmap) {
  LGate L;
  const size_t K = g.t.size(), D = 1ull << K;
  auto lq = [&](uint32_t q) -> uint32_t {
    if (q >= qmap.size() || qmap[q] < 0) throw std::logic_error("gate qubit outside the pass program");
    return (uint32_t)qmap[q];
  };
  std::vector<int> diag_in(K, 1);
  for (size_t i = 0; i < K; ++i)
    for (size_t r = 0; r < D && diag_in[i]; ++r)
      for (size_t c = 0; c < D; ++c)
        if ((((r ^ c) >> i) & 1u) && g.m[r * D + c] != cplx(0, 0)) { diag_in[i] = 0; break; }
  std::vector<size_t> ni, di;
  for (size_t i = 0; i < K; ++i) (diag_in[i] ? di : ni).push_back(i);
  const size_t nc = g.c.size();
  if (ni.empty()) {
    L.diag = true;
    for (size_t i = 0; i < K; ++i) L.Q.push_back(lq(g.t[i]));
    for (size_t i = 0; i < nc; ++i) L.Q.push_back(lq(g.c[i]));
    const size_t QD = 1ull << (K + nc);
    L.d.resize(QD);
    for (size_t x = 0; x < QD; ++x) {
      bool ok = true;
      for (size_t i = 0; i < nc; ++i)
        if (((x >> (K + i)) & 1u) != g.cval[i]) ok = false;
      const size_t tx = x & (D - 1);
      L.d[x] = ok ? g.m[tx * D + tx] : cplx(1, 0);
    }
    return L;
  }
  if (ni.size() > (size_t)kMaxDenseK) throw std::invalid_argument("gate with more than 5 non-diagonal targets");
  if (di.size() + nc > (size_t)kMaxSel) throw std::invalid_argument("gate with more than 6 diagonal/control qubits");
  for (auto i : ni) L.N.push_back(lq(g.t[i]));
  for (auto i : di) L.S.push_back(lq(g.t[i]));
  for (size_t i = 0; i < nc; ++i) L.S.push_back(lq(g.c[i]));
  const size_t nN = ni.size(), ND = 1ull << nN, nS = L.S.size(), SD = 1ull << nS;
  L.mats.assign(SD * ND * ND, cplx(0, 0));
  for (size_t s = 0; s < SD; ++s) {
    cplx* Ms = &L.mats[s * ND * ND];
    bool ctl_ok = true;
    for (size_t i = 0; i < nc; ++i)
      if (((s >> (di.size() + i)) & 1u) != g.cval[i]) ctl_ok = false;
    if (!ctl_ok) {
      for (size_t r = 0; r < ND; ++r) Ms[r * ND + r] = 1.0;
      L.ident |= 1ull << s;
      continue;
    }
    uint64_t dbits = 0;
    for (size_t j = 0; j < di.size(); ++j) dbits |= ((s >> j) & 1ull) << di[j];
    for (size_t rn = 0; rn < ND; ++rn)
      for (size_t cn = 0; cn < ND; ++cn) {
        uint64_t r = dbits, c = dbits;
        for (size_t j = 0; j < nN; ++j) {
          r |= ((rn >> j) & 1ull) << ni[j];
          c |= ((cn >> j) & 1ull) << ni[j];
        }
        Ms[rn * ND + cn] = g.m[r * D + c];
      }
    if (is_identity(Ms, ND)) L.ident |= 1ull << s;
  }
  return L;
}

struct PassPlan {
  uint32_t t, c;
  std::vector<uint32_t> E;
  std::vector<size_t> gates;
};

static int choose_c(const std::set<uint32_t>& R, uint32_t t, uint32_t cmin) {
  for (int c = (int)t; c >= (int)cmin; --c) {
    uint32_t cnt = 0;
    for (auto q : R)
      if (q >= (uint32_t)c) ++cnt;
    if ((uint32_t)c + cnt <= t) return c;
  }
  return -1;
}

struct Program {
  uint32_t L = 0;
  std::vector<PassArgs> passes;
  std::vector<Op> ops;
  std::vector<PhaseTerm> terms;
  std::vector<DiagTable> diags;
  std::vector<double> coefs;
  // phase tables of the wave passes' elementwise ops, built on the device
  // from the uploaded terms: (pass, op, tiles per group) per table
  struct PtabJob {
    size_t pass, op;
    uint64_t ntau, entries;
  };
  std::vector<PtabJob> ptab_jobs;
  uint64_t ptab_len = 0;  // complex entries
  DBuf d_args, d_ops, d_terms, d_diags, d_coefs, d_ptab;
  bool uploaded = false;

  // Elementwise ops of wave passes whose phase terms separate into the
  // factor tables (a slot term carries a lane or a base condition, not both;
  // a term on several slots carries neither) take the table path, while the
  // tables of the program fit kPhaseTableMaxEntries.  QOMPSIM_PHASE_TABLES=0
  // keeps every op on the per-term loop.
  void plan_phase_tables() {
    ptab_jobs.clear();
    ptab_len = 0;
    for (auto& op : ops) op.ptab_off = kNoPhaseTable;
    const char* e = std::getenv("QOMPSIM_PHASE_TABLES");
    if (e && e[0] == '0') return;
    for (size_t pi = 0; pi < passes.size(); ++pi) {
      const PassArgs& a = passes[pi];
      if (a.engine != ENGINE_WAVE || a.t < 7 || a.t > 10 || a.c != 6 || L < a.t) continue;
      if (L - a.t > kPhaseTableMaxTauLog) continue;
      const uint64_t ntau = 1ull << (L - a.t);
      for (uint32_t o = a.op_begin; o < a.op_end; ++o) {
        Op& op = ops[o];
        if (op.kind != OP_ELEM || op.nterms < kPhaseTableMinTerms) continue;
        uint32_t slots = 0, flags = 0;
        bool ok = true;
        // the base bits tested by slot-free parts with lane and base
        // conditions: few of them key a small lane table K
        uint64_t kmask = 0;
        for (uint32_t i = 0; i < op.nterms; ++i) {
          const PhaseTerm& T = terms[op.term_off + i];
          if (T.bval & ~T.bmask) continue;  // padding: never matches
          const uint32_t lm = T.tmask & 63u, rm = (T.tmask >> 6) & 15u, rv = (T.tval >> 6) & 15u;
          const bool has_free = rm == 0u || (!(rm & (rm - 1u)) && rv == 0u);
          if (has_free && lm && T.bmask) kmask |= T.bmask;
        }
        const bool keyed = (uint32_t)__builtin_popcountll(kmask) <= kPtabMaxKeyBits;
        // otherwise terms on one lane bit and base bits: the lane table M
        // while it fits, else the lane-bit angles G
        const bool use_g = !keyed && L - a.t > kPhaseTableMaxTauLogM;
        auto slot_free = [&](uint32_t lm, uint32_t lv, uint64_t bm) {
          if (keyed) {
            flags |= (lm && bm) ? kPtabK : (bm ? kPtabU : kPtabQ);
          } else if (use_g && lm && bm && !(lm & (lm - 1u))) {
            slots |= lm << 8;
            flags |= kPtabG;
            if (!lv) flags |= kPtabU;
          } else {
            flags |= (lm && bm) ? kPtabM : (bm ? kPtabU : kPtabQ);
          }
        };
        for (uint32_t i = 0; i < op.nterms; ++i) {
          const PhaseTerm& T = terms[op.term_off + i];
          if (T.bval & ~T.bmask) continue;  // padding: never matches
          const uint32_t lm = T.tmask & 63u, lv = T.tval & 63u, rm = (T.tmask >> 6) & 15u, rv = (T.tval >> 6) & 15u;
          if (rm == 0u) {
            slot_free(lm, lv, T.bmask);
          } else if (rm & (rm - 1u)) {
            ok = ok && lm == 0u && T.bmask == 0ull;
            flags |= kPtabRows;
          } else {
            ok = ok && !(lm != 0u && T.bmask != 0ull);
            slots |= rm;
            if (T.bmask) flags |= kPtabFt;
            if (rv == 0u) slot_free(lm, lv, T.bmask);
          }
        }
        if ((flags & kPtabM) && L - a.t > kPhaseTableMaxTauLogM) ok = false;
        // a lane table M takes every slot-free part (one read per tile)
        if (flags & kPtabM) flags &= ~(kPtabQ | kPtabU);
        // only the tables that hold a term take room
        const uint32_t nk = (flags & kPtabK) ? (uint32_t)__builtin_popcountll(kmask) : 0u;
        const uint64_t q = kPtabF + ((flags & kPtabFt) ? 4ull * ntau : 0ull);
        const uint64_t g = q + 64 + ((flags & kPtabU) ? ntau : 0ull);
        const uint64_t m = g + ((flags & kPtabG) ? 8ull * ntau : 0ull);
        const uint64_t len = m + ((flags & kPtabK) ? (64ull << nk) : (flags & kPtabM) ? 64ull * ntau : 0ull);
        if (!ok || ptab_len + len > kPhaseTableMaxEntries) continue;
        op.ptab_off = ptab_len;
        op.ptab_q = (uint32_t)q;
        op.ptab_g = (uint32_t)g;
        op.ptab_m = (uint32_t)m;
        op.ptab_slots = slots;
        op.ptab_flags = flags;
        op.ptab_nk = nk;
        for (uint32_t i = 0, b = 0; b < 64 && i < nk; ++b)
          if ((kmask >> b) & 1ull) op.ptab_kpos[i++] = (uint8_t)b;
        ptab_jobs.push_back({pi, o, ntau, len});
        ptab_len += (len + 63) & ~63ull;
      }
    }
  }

  void upload_all() {
    plan_phase_tables();
    if (ptab_len) d_ptab.ensure(ptab_len * 2 * sizeof(double));
    upload(d_ops, ops.data(), ops.size() * sizeof(Op));
    // kTermGroup never-matching terms past the end: the kernels read the
    // next group of an op's terms ahead of the current one
    std::vector<PhaseTerm> tpad(terms);
    for (uint32_t i = 0; i < kTermGroup; ++i) {
      PhaseTerm T{};
      T.bval = 1;
      tpad.push_back(T);
    }
    upload(d_terms, tpad.data(), tpad.size() * sizeof(PhaseTerm));
    upload(d_diags, diags.data(), diags.size() * sizeof(DiagTable));
    upload(d_coefs, coefs.data(), coefs.size() * sizeof(double));
    for (auto& p : passes) {
      p.ops = d_ops.as<Op>();
      p.terms = d_terms.as<PhaseTerm>();
      p.diags = d_diags.as<DiagTable>();
      p.coefs = d_coefs.as<double>();
      p.ptab = ptab_len ? d_ptab.as<double>() : nullptr;
    }
    upload(d_args, passes.data(), passes.size() * sizeof(PassArgs));
    if (std::getenv("QOMPSIM_PLAN_LOG"))
      fprintf(stderr, "[plan] phase tables: ops=%zu entries=%llu\n", ptab_jobs.size(), (unsigned long long)ptab_len);
    for (const auto& j : ptab_jobs)
      HIPCHECK(launch_phase_table(dargs(j.pass), d_ops.as<Op>() + j.op, j.ntau, j.entries, d_ptab.as<double>(),
                                  stream()));
    HIPCHECK(hipStreamSynchronize(stream()));
    uploaded = true;
  }
  const PassArgs* dargs(size_t i) const { return d_args.as<PassArgs>() + i; }
  // forget the plan (device buffers are kept and reused by the next upload)
  void reset() {
    L = 0;
    passes.clear();
    ops.clear();
    terms.clear();
    diags.clear();
    coefs.clear();
    ptab_jobs.clear();
    ptab_len = 0;
    uploaded = false;
  }
};

static void emit_pass(Program& P, const PassPlan& pp, const std::vector<LGate>& lg) {
  // dense gates with real matrices, or one-target matrices with real diagonal
  // and imaginary off-diagonal entries, take the kernels' cheaper paths
  // (QOMPSIM_REAL_DENSE=0 sends them through the complex ones)
  const char* er = std::getenv("QOMPSIM_REAL_DENSE");
  const bool real_dense = !(er && er[0] == '0');
  PassArgs a{};
  a.t = pp.t;
  a.c = pp.c;
  a.nE = (uint32_t)pp.E.size();
  a.L = P.L;
  for (size_t i = 0; i < pp.E.size(); ++i) a.E[i] = pp.E[i];
  a.op_begin = (uint32_t)P.ops.size();
  auto tpos = [&](uint32_t q) -> int {
    if (q < pp.c) return (int)q;
    for (size_t i = 0; i < pp.E.size(); ++i)
      if (pp.E[i] == q) return (int)(pp.c + i);
    return -1;
  };
  bool elem_open = false;
  // the open elementwise op is the last op; its terms end the term list
  auto close_elem = [&]() {
    if (!elem_open) return;
    Op& op = P.ops.back();
    while (op.nterms % kTermGroup) {
      PhaseTerm T{};
      T.bval = 1;  // never matches: (base & 0) != 1
      P.terms.push_back(T);
      op.nterms++;
    }
    elem_open = false;
  };
  // A diagonal gate moves past later dense gates whose non-diagonal targets
  // avoid its qubits (the two commute), so that the pass's diagonal gates
  // gather into fewer elementwise ops, each of which evaluates its phase
  // factors once (RY RZ RY RZ ... on different qubits: one op for all RZ).
  // A conflicting dense gate sends the whole pending run ahead of itself:
  // sending only the gates it conflicts with made one op per RY of HEA's
  // next layer.
  // A gradient point Re<lambda|iP psi> is read where psi and lambda have been
  // carried through the same gates, so a diagonal gate D moves past it when D
  // commutes with P, i.e. avoids P's X/Y qubits.  QOMPSIM_DIAG_GATHER=0 keeps
  // the order.
  std::vector<size_t> order;
  order.reserve(pp.gates.size());
  {
    const char* eg = std::getenv("QOMPSIM_DIAG_GATHER");
    const bool gather = !(eg && eg[0] == '0');
    std::vector<size_t> run;
    uint64_t run_q = 0;
    auto flush = [&]() {
      order.insert(order.end(), run.begin(), run.end());
      run.clear();
      run_q = 0;
    };
    for (size_t gi : pp.gates) {
      const LGate& g = lg[gi];
      if (!gather) {
        order.push_back(gi);
        continue;
      }
      if (g.grad >= 0) {
        if (g.gx & run_q) flush();
        order.push_back(gi);
        continue;
      }
      if (g.diag) {
        run.push_back(gi);
        for (auto q : g.Q) run_q |= 1ull << q;
        continue;
      }
      uint64_t nq = 0;
      for (auto q : g.N) nq |= 1ull << q;
      if (nq & run_q) flush();
      order.push_back(gi);
    }
    flush();
  }
  for (size_t gi : order) {
    const LGate& g = lg[gi];
    if (g.noop) continue;
    if (g.grad >= 0) {
      close_elem();
      Op op{};
      op.kind = OP_GRAD;
      op.gidx = (uint32_t)g.grad;  // parameter index; renumbered per pass by the caller
      op.nY = g.nY;
      for (uint32_t q = 0; q < 64; ++q) {
        const uint64_t bit = 1ull << q;
        if (!((g.gx | g.gz) & bit)) continue;
        const int p = tpos(q);
        if (g.gx & bit) {
          if (p < 0) throw std::logic_error("planner: gradient flip outside the tile");
          op.gfx |= 1ull << p;
        }
        if (g.gz & bit) {
          if (p >= 0) op.gzt |= 1ull << p;
          else op.gzb |= bit;
        }
      }
      P.ops.push_back(op);
      continue;
    }
    if (g.diag) {
      if (!elem_open) {
        Op op{};
        op.kind = OP_ELEM;
        op.term_off = (uint32_t)P.terms.size();
        op.diag_off = (uint32_t)P.diags.size();
        P.ops.push_back(op);
        elem_open = true;
      }
      Op& op = P.ops.back();
      bool unit = true;
      for (auto& v : g.d)
        if (std::fabs(std::norm(v) - 1.0) > 1e-13) unit = false;
      if (unit) {
        for (size_t x = 0; x < g.d.size(); ++x) {
          const double ang = std::atan2(g.d[x].imag(), g.d[x].real());
          if (ang == 0.0) continue;
          PhaseTerm T{};
          for (size_t j = 0; j < g.Q.size(); ++j) {
            const uint32_t q = g.Q[j];
            const uint64_t bit = (x >> j) & 1ull;
            const int p = tpos(q);
            if (p >= 0) {
              T.tmask |= 1u << p;
              T.tval |= (uint32_t)bit << p;
            } else {
              T.bmask |= 1ull << q;
              T.bval |= bit << q;
            }
          }
          T.angle = ang;
          P.terms.push_back(T);
          op.nterms++;
        }
      } else {
        if (g.Q.size() > 8) throw std::invalid_argument("non-unitary diagonal gate on more than 8 qubits");
        DiagTable Dt{};
        Dt.m = (uint32_t)g.Q.size();
        Dt.coef_off = (uint32_t)(P.coefs.size() / 2);
        for (size_t j = 0; j < g.Q.size(); ++j) {
          const int p = tpos(g.Q[j]);
          Dt.tile[j] = p >= 0 ? 1 : 0;
          Dt.pos[j] = (uint16_t)(p >= 0 ? p : g.Q[j]);
        }
        for (auto& v : g.d) { P.coefs.push_back(v.real()); P.coefs.push_back(v.imag()); }
        P.diags.push_back(Dt);
        op.ndiag++;
      }
      continue;
    }
    close_elem();
    Op op{};
    op.kind = OP_DENSE;
    op.k = (uint32_t)g.N.size();
    op.nsel = (uint32_t)g.S.size();
    op.ident_mask = g.ident;
    std::vector<uint16_t> srt;
    for (size_t j = 0; j < g.N.size(); ++j) {
      const int p = tpos(g.N[j]);
      if (p < 0) throw std::logic_error("planner: dense target outside the tile");
      op.tpos[j] = (uint16_t)p;
      srt.push_back((uint16_t)p);
    }
    std::sort(srt.begin(), srt.end());
    for (size_t j = 0; j < srt.size(); ++j) op.tsorted[j] = srt[j];
    for (size_t j = 0; j < g.S.size(); ++j) {
      const int p = tpos(g.S[j]);
      op.sel_tile[j] = p >= 0 ? 1 : 0;
      op.sel_pos[j] = (uint16_t)(p >= 0 ? p : g.S[j]);
    }
    op.mat_off = (uint32_t)(P.coefs.size() / 2);
    bool all_real = true, rdio = op.k == 1;
    for (size_t i = 0; i < g.mats.size(); ++i) {
      const cplx v = g.mats[i];
      P.coefs.push_back(v.real());
      P.coefs.push_back(v.imag());
      if (v.imag() != 0.0) all_real = false;
      const bool on_diag = (i & 3u) == 0 || (i & 3u) == 3;  // 2x2 entries 0 and 3
      if (on_diag ? v.imag() != 0.0 : v.real() != 0.0) rdio = false;
    }
    op.real = !real_dense ? 0u : all_real ? kDenseReal : rdio ? kDenseRealDiagImagOff : 0u;
    P.ops.push_back(op);
    a.maxk = std::max(a.maxk, op.k);
  }
  close_elem();
  a.op_end = (uint32_t)P.ops.size();
  P.passes.push_back(a);
}

// Greedy split of a gate list into LDS passes: a pass takes gates in order
// while the union of their non-diagonal targets fits into a tile of 2^t
// amplitudes whose low c >= cmin qubits are contiguous.  A gate that cannot
// meet cmin on its own gets a pass with a shorter contiguous run.  Diagonal
// gates fit every pass.  With codec_ends the first and last pass keep
// c >= kChunkAmpsLog2 (codec chunks are 32 contiguous amplitudes); an
// op-free pass is added at either end when needed.
static void plan_program_engine(Program& P, const std::vector<LGate>& lg, uint32_t L, uint32_t tmax, uint32_t cmin,
                                bool at_least_one_pass, bool codec_ends, bool wave, uint32_t wave_bits = kWaveTile,
                                std::vector<std::vector<size_t>>* pass_gates = nullptr) {
  P.L = L;
  const uint32_t t = wave ? wave_bits : std::min(tmax, L);
  if (wave) cmin = 6;
  cmin = std::min(cmin, t);
  // Commutation-aware greedy scheduling: a pass scans the remaining gates in
  // order; a gate joins when it commutes with every gate deferred before it
  // (no shared qubit on which either acts non-diagonally) and its
  // non-diagonal targets fit the tile; otherwise it is deferred to a later
  // pass.  Applied gates keep their relative order.
  std::vector<PassPlan> plans;
  auto qmask = [](const std::vector<uint32_t>& v) {
    uint64_t m = 0;
    for (auto q : v) m |= 1ull << q;
    return m;
  };
  std::vector<uint64_t> ndm(lg.size()), allm(lg.size());
  for (size_t i = 0; i < lg.size(); ++i) {
    const LGate& g = lg[i];
    ndm[i] = g.diag ? 0ull : qmask(g.N);
    allm[i] = g.diag ? qmask(g.Q) : (qmask(g.N) | qmask(g.S));
    if (!g.diag) {
      std::set<uint32_t> Rg(g.N.begin(), g.N.end());
      if (choose_c(Rg, t, 0) < 0) throw std::invalid_argument("gate does not fit in one LDS tile");
    }
  }
  std::vector<size_t> remaining(lg.size());
  std::iota(remaining.begin(), remaining.end(), 0);
  while (!remaining.empty()) {
    std::set<uint32_t> R;
    PassPlan cur{t, t, {}, {}};
    std::vector<size_t> deferred;
    uint64_t d_all = 0, d_nd = 0;
    for (size_t i : remaining) {
      const LGate& g = lg[i];
      bool take = false;
      if (!((ndm[i] & d_all) || (allm[i] & d_nd))) {
        if (g.diag) {
          take = true;
        } else {
          std::set<uint32_t> R2 = R, Rg(g.N.begin(), g.N.end());
          R2.insert(g.N.begin(), g.N.end());
          if (choose_c(R2, t, cmin) >= 0) take = true;
          else if (choose_c(Rg, t, cmin) < 0 && choose_c(R2, t, 0) >= 0) take = true;  // needs a short run anyway
          if (take) R = R2;
        }
      }
      if (take) {
        cur.gates.push_back(i);
      } else {
        deferred.push_back(i);
        d_all |= allm[i];
        d_nd |= ndm[i];
      }
    }
    if (cur.gates.empty()) throw std::logic_error("planner: no progress");
    int c = choose_c(R, t, cmin);
    if (c < 0) c = choose_c(R, t, 0);
    if (c < 0) throw std::logic_error("planner: tile overflow");
    cur.c = (uint32_t)c;
    for (auto q : R)
      if (q >= (uint32_t)c) cur.E.push_back(q);
    plans.push_back(cur);
    remaining.swap(deferred);
  }
  if (plans.empty() && at_least_one_pass) plans.push_back(PassPlan{t, t, {}, {}});
  if (wave) {
    // lane bits are local qubits 0..5; every other tile qubit is a register slot
    for (auto& pp : plans) {
      std::vector<uint32_t> E;
      for (uint32_t q = 6; q < pp.c; ++q) E.push_back(q);
      E.insert(E.end(), pp.E.begin(), pp.E.end());
      std::sort(E.begin(), E.end());
      if (E.size() != wave_bits - 6) throw std::logic_error("planner: wave tile register qubit count");
      pp.c = 6;
      pp.E = E;
    }
  }
  if (codec_ends) {
    const uint32_t need = std::min<uint32_t>((uint32_t)kChunkAmpsLog2, t);
    if (plans.front().c < need) plans.insert(plans.begin(), PassPlan{t, t, {}, {}});
    if (plans.back().c < need) plans.push_back(PassPlan{t, t, {}, {}});
  }
  for (auto& pp : plans) {
    emit_pass(P, pp, lg);
    P.passes.back().engine = wave ? ENGINE_WAVE : ENGINE_LDS;
    if (pass_gates) pass_gates->push_back(pp.gates);
  }
}

// The wave-resident engine takes programs on at least 10 local qubits whose
// gates have at most two non-diagonal targets; the LDS-tile engine takes the
// rest.  QOMPSIM_ENGINE=lds forces the LDS-tile engine.
static bool wave_engine_ok(const std::vector<LGate>& lg, uint32_t L) {
  const char* e = std::getenv("QOMPSIM_ENGINE");
  if (e && std::string(e) == "lds") return false;
  if (L < kWaveTile) return false;
  for (auto& g : lg)
    if (!g.diag && g.N.size() > 2) return false;
  return true;
}

// Compressed programs (codec_ends) that fit one wave pass of 1024-amplitude
// tiles are re-planned on 512-amplitude tiles (eight rows per lane, three
// waves per SIMD, more latency hidden in the codec) when that still takes one
// pass.  QOMPSIM_WAVE_TILE_BITS=10 keeps the larger tile.
static uint32_t compressed_wave_bits() {
  const char* e = std::getenv("QOMPSIM_WAVE_TILE_BITS");
  return (e && std::atoi(e) == (int)kWaveTile) ? kWaveTile : kWaveTileSmall;
}

// An exact SWAP: two non-diagonal targets, no selector, the 4x4 permutation.
static bool is_swap_gate(const LGate& g) {
  if (g.grad >= 0 || g.diag || g.N.size() != 2 || !g.S.empty() || g.mats.size() != 16) return false;
  static const int perm[4] = {0, 2, 1, 3};
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c)
      if (g.mats[r * 4 + c] != cplx(c == perm[r] ? 1.0 : 0.0, 0.0)) return false;
  return true;
}

// LDS slots of the encoder's layout transpose for a tile whose bits move by
// the swaps `ts` (each a pair of tile qubits, pairwise disjoint): the value
// at tile position p = 64 v + l goes to pi(p) = 64 v' + l', whose slot is
// RS v' + l' + (l' >> PS) (the padded layout of the wave kernel).  pi(p)
// splits into a part from l and a part from v with disjoint bits, so the
// slot is tperm_lane[l] + tperm_row[v].
static void set_tile_permutation(PassArgs& a, const std::vector<std::pair<uint32_t, uint32_t>>& ts) {
  const uint32_t RB = a.t - 6, KV = 1u << RB;
  const uint32_t PS = KV == 16 ? 4u : 3u, RS = 64u + (64u >> PS);
  uint32_t perm[16];
  for (uint32_t i = 0; i < a.t; ++i) perm[i] = i;
  auto tbit = [&](uint32_t q) -> uint32_t {
    if (q < 6) return q;
    for (uint32_t j = 0; j < a.nE; ++j)
      if (a.E[j] == q) return 6 + j;
    throw std::logic_error("planner: swapped qubit outside the tile");
  };
  for (auto& pr : ts) std::swap(perm[tbit(pr.first)], perm[tbit(pr.second)]);
  auto slot = [&](uint32_t p) { const uint32_t l = p & 63u, v = p >> 6; return RS * v + l + (l >> PS); };
  for (uint32_t l = 0; l < 64; ++l) {
    uint32_t p = 0;
    for (uint32_t i = 0; i < 6; ++i) p |= ((l >> i) & 1u) << perm[i];
    a.tperm_lane[l] = (uint16_t)slot(p);
  }
  for (uint32_t v = 0; v < KV; ++v) {
    uint32_t p = 0;
    for (uint32_t j = 0; j < RB; ++j) p |= ((v >> j) & 1u) << perm[6 + j];
    a.tperm_row[v] = (uint16_t)slot(p);
  }
  a.tperm = 1;
}

// M (4x4, index bit 0 <-> the first target) as (A (x) B) diag(d): A on the
// second target, B on the first, each column c of M being d_c times
// A[:, c >> 1] (x) B[:, c & 1].  The columns of A and B are read off M's
// columns, scaled so that their diagonal entries are real and non-negative, d
// follows by projection, and the product must give M back to 1e-14.
static bool factor_product(const cplx* M, cplx* A, cplx* Bm, cplx* d) {
  auto K = [&](int c, int r1, int r0) { return M[(2 * r1 + r0) * 4 + c]; };
  cplx a[2][2], b[2][2];
  for (int c1 = 0; c1 < 2; ++c1) {
    const int c = 2 * c1;
    const int r0 = std::norm(K(c, 0, 1)) + std::norm(K(c, 1, 1)) > std::norm(K(c, 0, 0)) + std::norm(K(c, 1, 0)) ? 1 : 0;
    const double nn = std::sqrt(std::norm(K(c, 0, r0)) + std::norm(K(c, 1, r0)));
    if (!(nn > 1e-150)) return false;
    a[c1][0] = K(c, 0, r0) / nn;
    a[c1][1] = K(c, 1, r0) / nn;
  }
  for (int c0 = 0; c0 < 2; ++c0) {
    const int c = c0;
    const int r1 = std::norm(K(c, 1, 0)) + std::norm(K(c, 1, 1)) > std::norm(K(c, 0, 0)) + std::norm(K(c, 0, 1)) ? 1 : 0;
    const double nn = std::sqrt(std::norm(K(c, r1, 0)) + std::norm(K(c, r1, 1)));
    if (!(nn > 1e-150)) return false;
    b[c0][0] = K(c, r1, 0) / nn;
    b[c0][1] = K(c, r1, 1) / nn;
  }
  for (int k = 0; k < 2; ++k) {
    if (std::abs(a[k][k]) > 1e-12) {
      const cplx ph = std::conj(a[k][k]) / std::abs(a[k][k]);
      a[k][0] *= ph;
      a[k][1] *= ph;
    }
    if (std::abs(b[k][k]) > 1e-12) {
      const cplx ph = std::conj(b[k][k]) / std::abs(b[k][k]);
      b[k][0] *= ph;
      b[k][1] *= ph;
    }
  }
  for (int c = 0; c < 4; ++c) {
    const int c1 = c >> 1, c0 = c & 1;
    cplx s = 0.0;
    for (int r1 = 0; r1 < 2; ++r1)
      for (int r0 = 0; r0 < 2; ++r0) s += std::conj(a[c1][r1] * b[c0][r0]) * K(c, r1, r0);
    d[c] = s;
    for (int r1 = 0; r1 < 2; ++r1)
      for (int r0 = 0; r0 < 2; ++r0)
        if (std::abs(K(c, r1, r0) - s * a[c1][r1] * b[c0][r0]) > 1e-14) return false;
  }
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 2; ++c) {
      A[r * 2 + c] = a[c][r];
      Bm[r * 2 + c] = b[c][r];
    }
  return true;
}

// A two-target dense gate whose matrix is a product of one-qubit matrices and
// a diagonal, M = (A (x) B) D or M = D (A (x) B), becomes the diagonal and the
// two one-qubit gates (QuantumCircuitOptimizer merges RZZ with the RX gates on
// its qubits into such matrices): two one-target products and a phase op
// instead of a 4x4 complex product per four amplitudes, and the one-qubit
// gates may take different passes.  Likewise a one-target gate whose matrices
// for the selector values differ by a diagonal factor (the optimizer's H(j)
// merged with CP(j, k): H on j selected by k) becomes that diagonal and one
// matrix without selectors, which may be real.  QOMPSIM_SPLIT_PRODUCTS=0 keeps
// the merged matrices.
static std::vector<LGate> split_product_gates(const std::vector<LGate>& lg) {
  const char* e = std::getenv("QOMPSIM_SPLIT_PRODUCTS");
  if (e && *e == '0') return lg;
  std::vector<LGate> out;
  out.reserve(lg.size());
  for (const LGate& g : lg) {
    if (g.grad < 0 && !g.noop && !g.diag && g.N.size() == 1 && !g.S.empty() &&
        g.mats.size() == ((size_t)4 << g.S.size())) {
      // one target with selectors: M_s = B D_s (or D_s B) for every selector
      // value s, with B = M_0 and D_s diagonal, is the diagonal over the
      // target and the selectors, then B without selectors
      const size_t ns = g.mats.size() / 4;
      const cplx* B0 = g.mats.data();
      const cplx det = B0[0] * B0[3] - B0[1] * B0[2];
      bool done = false;
      if (std::abs(det) > 1e-12) {
        const cplx Bi[4] = {B0[3] / det, -B0[1] / det, -B0[2] / det, B0[0] / det};
        for (int after = 0; after < 2 && !done; ++after) {
          std::vector<cplx> dd(2 * ns);
          bool ok = true;
          for (size_t sv = 0; sv < ns && ok; ++sv) {
            const cplx* Ms = &g.mats[4 * sv];
            cplx Ds[4];
            for (int r = 0; r < 2; ++r)
              for (int c = 0; c < 2; ++c)
                Ds[r * 2 + c] = after ? Ms[r * 2] * Bi[c] + Ms[r * 2 + 1] * Bi[2 + c]   // M_s B^-1
                                      : Bi[r * 2] * Ms[c] + Bi[r * 2 + 1] * Ms[2 + c];  // B^-1 M_s
            if (std::abs(Ds[1]) > 1e-14 || std::abs(Ds[2]) > 1e-14) ok = false;
            dd[2 * sv] = Ds[0];
            dd[2 * sv + 1] = Ds[3];
          }
          if (!ok) continue;
          // the product must give every M_s back
          for (size_t sv = 0; sv < ns && ok; ++sv)
            for (int r = 0; r < 2; ++r)
              for (int c = 0; c < 2; ++c) {
                const cplx v = after ? dd[2 * sv + r] * B0[r * 2 + c] : B0[r * 2 + c] * dd[2 * sv + c];
                if (std::abs(v - g.mats[4 * sv + r * 2 + c]) > 1e-14) ok = false;
              }
          if (!ok) continue;
          LGate D;
          D.diag = true;
          D.Q.push_back(g.N[0]);
          D.Q.insert(D.Q.end(), g.S.begin(), g.S.end());
          D.d = dd;  // index bit 0 <-> the target, bit 1 + j <-> selector j
          LGate Bg;
          Bg.N = g.N;
          Bg.mats.assign(B0, B0 + 4);
          if (is_identity(B0, 2)) Bg.ident = 1;
          if (!after) out.push_back(D);
          out.push_back(Bg);
          if (after) out.push_back(D);
          done = true;
        }
      }
      if (!done) out.push_back(g);
      continue;
    }
    if (g.grad >= 0 || g.noop || g.diag || g.N.size() != 2 || !g.S.empty() || g.mats.size() != 16 ||
        is_swap_gate(g)) {
      out.push_back(g);
      continue;
    }
    cplx A[4], Bm[4], dg[4];
    bool diag_first = factor_product(g.mats.data(), A, Bm, dg);
    if (!diag_first) {
      cplx Mt[16];
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) Mt[c * 4 + r] = g.mats[r * 4 + c];
      cplx At[4], Bt[4];
      if (!factor_product(Mt, At, Bt, dg)) {
        out.push_back(g);
        continue;
      }
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) {
          A[r * 2 + c] = At[c * 2 + r];
          Bm[r * 2 + c] = Bt[c * 2 + r];
        }
    }
    LGate D;
    D.diag = true;
    D.Q = {g.N[0], g.N[1]};
    D.d.assign(dg, dg + 4);
    auto one = [&](uint32_t q, const cplx* m) {
      LGate G;
      G.N = {q};
      G.mats.assign(m, m + 4);
      if (is_identity(m, 2)) G.ident = 1;
      return G;
    };
    if (diag_first) out.push_back(D);
    out.push_back(one(g.N[1], A));
    out.push_back(one(g.N[0], Bm));
    if (!diag_first) out.push_back(D);
  }
  return out;
}

static void plan_program(Program& P, const std::vector<LGate>& lg0, uint32_t L, uint32_t tmax, uint32_t cmin,
                         bool at_least_one_pass, bool codec_ends = false) {
  // Trailing SWAP gates of a compressed program move data without
  // arithmetic (wave engine; QOMPSIM_STORE_SWAPS=0 keeps them as gates):
  // - on qubits at or above the chunk bits they move whole chunks and become
  //   the last (encoding) pass's chunk permutation;
  // - with a qubit below the chunk bits, when both qubits are tile bits of a
  //   single-pass program, they become a permutation of the tile in the
  //   encoder's layout transpose.  Such swaps are disjoint from every other
  //   folded swap, so the order of the two permutations does not matter.
  std::vector<LGate> lg = split_product_gates(lg0);
  std::vector<std::pair<uint32_t, uint32_t>> swaps, tswaps;
  std::vector<LGate> tswap_gates;
  const char* es = std::getenv("QOMPSIM_STORE_SWAPS");
  if (codec_ends && wave_engine_ok(lg0, L) && !(es && *es == '0')) {
    uint64_t tq = 0, allq = 0;
    while (!lg.empty() && is_swap_gate(lg.back())) {
      const uint32_t qa = lg.back().N[0], qb = lg.back().N[1];
      const uint64_t m = (1ull << qa) | (1ull << qb);
      if (qa >= (uint32_t)kChunkAmpsLog2 && qb >= (uint32_t)kChunkAmpsLog2) {
        if (swaps.size() >= (size_t)kMaxStoreSwaps || (m & tq)) break;
        swaps.emplace_back(qa, qb);
      } else {
        if (tswaps.size() >= 6 || (m & allq)) break;
        tswaps.emplace_back(qa, qb);
        tswap_gates.push_back(lg.back());
        tq |= m;
      }
      allq |= m;
      lg.pop_back();
    }
    std::reverse(swaps.begin(), swaps.end());
    std::reverse(tswaps.begin(), tswaps.end());
    std::reverse(tswap_gates.begin(), tswap_gates.end());
  }
  const bool wave = wave_engine_ok(lg0, L);
  auto plan_all = [&](const std::vector<LGate>& G, bool need_pass) {
    plan_program_engine(P, G, L, tmax, cmin, at_least_one_pass || need_pass, codec_ends, wave);
    if (wave && codec_ends && P.passes.size() == 1 && compressed_wave_bits() == kWaveTileSmall) {
      Program Q;
      plan_program_engine(Q, G, L, tmax, cmin, at_least_one_pass || need_pass, codec_ends, true, kWaveTileSmall);
      if (Q.passes.size() == 1) {
        P.reset();
        plan_program_engine(P, G, L, tmax, cmin, at_least_one_pass || need_pass, codec_ends, true, kWaveTileSmall);
      }
    }
  };
  if (!tswaps.empty()) {
    // the tile swaps reserve their qubits in the tile as op-free gates
    std::vector<LGate> lgp = lg;
    for (const auto& g : tswap_gates) {
      lgp.push_back(g);
      lgp.back().noop = true;
    }
    plan_all(lgp, true);
    if (P.passes.size() == 1 && P.passes[0].engine == ENGINE_WAVE) {
      set_tile_permutation(P.passes[0], tswaps);
    } else {
      // more than one pass: the tile swaps stay gates
      P.reset();
      lg.insert(lg.end(), tswap_gates.begin(), tswap_gates.end());
      tswaps.clear();
      plan_all(lg, !swaps.empty());
    }
  } else {
    plan_all(lg, !swaps.empty());
  }
  if (!swaps.empty()) {
    PassArgs& last = P.passes.back();
    last.nswap = (uint32_t)swaps.size();
    for (size_t i = 0; i < swaps.size(); ++i) {
      last.swp[2 * i] = (uint8_t)swaps[i].first;
      last.swp[2 * i + 1] = (uint8_t)swaps[i].second;
    }
  }
  if (std::getenv("QOMPSIM_PLAN_LOG")) {
    fprintf(stderr, "[plan] L=%u gates=%zu passes=%zu\n", L, lg.size(), P.passes.size());
    for (auto& a : P.passes) {
      uint32_t nd = 0, ne = 0, nt = 0;
      for (uint32_t o = a.op_begin; o < a.op_end; ++o) {
        if (P.ops[o].kind == OP_DENSE) ++nd;
        else { ++ne; nt += P.ops[o].nterms; }
      }
      fprintf(stderr, "[plan]   engine=%s t=%u c=%u E=", a.engine == ENGINE_WAVE ? "wave" : "lds", a.t, a.c);
      for (uint32_t i = 0; i < a.nE; ++i) fprintf(stderr, "%u%s", a.E[i], i + 1 < a.nE ? "," : "");
      fprintf(stderr, " dense=%u elem=%u terms=%u maxk=%u swaps=%u tile_perm=%u\n", nd, ne, nt, a.maxk, a.nswap,
              a.tperm);
    }
  }
}

// Launch pass i of a program over `groups` group slots of 2^L amplitudes.
static hipError_t launch_program_pass(const Program& P, size_t i, int in_mode, int out_mode, const PassDyn& dy,
                                      uint64_t groups, hipStream_t st) {
  const PassArgs& a = P.passes[i];
  if ((a.nswap || a.tperm) && (a.engine != ENGINE_WAVE || out_mode != OUT_ENC)) return hipErrorInvalidValue;
  if (a.engine == ENGINE_WAVE) {
    if (a.t != kWaveTile && a.t != kWaveTileSmall) return hipErrorInvalidValue;
    return launch_wave((int)a.t, in_mode, out_mode, P.dargs(i), dy, groups << (P.L - a.t), st);
  }
  return launch_pass(in_mode, out_mode, (int)a.t, a.maxk > 3, P.dargs(i), dy, groups << (P.L - a.t), st);
}

// --------------------------------------------------------- reductions (DD)
struct DDAcc {
  double s = 0.0, c = 0.0;
  void add(double x) {
    const double t = s + x;
    const double bp = t - s;
    c += (s - (t - bp)) + (x - bp);
    s = t;
  }
  void merge(double s2, double c2) { add(s2); c += c2; }
  double value() const { return s + c; }
};

struct Partials {
  DBuf d;
  std::vector<double> h;
  void ensure(size_t nv, unsigned grid) {
    d.ensure(sizeof(double) * 2 * nv * grid);
    h.resize(2 * nv * grid);
  }
  void fetch(size_t nv, unsigned grid) {
    HIPCHECK(hipMemcpyAsync(h.data(), d.p, sizeof(double) * 2 * nv * grid, hipMemcpyDeviceToHost, stream()));
    HIPCHECK(hipStreamSynchronize(stream()));
  }
  void accumulate(size_t nv, unsigned grid, DDAcc* acc) const {
    for (unsigned b = 0; b < grid; ++b)
      for (size_t v = 0; v < nv; ++v) acc[v].merge(h[(b * nv + v) * 2], h[(b * nv + v) * 2 + 1]);
  }
};

}  // namespace qsim

// ====================================================================== SV
using namespace qsim;

struct qsim_sv {
  uint32_t n = 0;
  uint64_t dim = 0;
  DBuf d;
  mutable Partials part;
  // adjoint work buffers (kept between calls): lambda, the uncomputed psi,
  // per-tile gradient partials, reduced points, Pauli-sum tables
  DBuf adj_lam, adj_psi, adj_part, adj_out, adj_stage, adj_groups, adj_terms;
  // the lambda-only engine's copies of psi at reverse pass starts
  std::vector<std::unique_ptr<DBuf>> adj_ck;
  uint64_t adj_ck_bytes() const {
    uint64_t b = 0;
    for (auto& c : adj_ck) b += c->bytes;
    return b;
  }
  double2* ptr() const { return reinterpret_cast<double2*>(d.p); }
};

namespace qsim {
static uint32_t default_tile_bits() {
  const char* e = std::getenv("QOMPSIM_TILE_BITS");
  if (e) {
    int v = std::atoi(e);
    if (v >= 5 && v <= kMaxTileBits) return (uint32_t)v;
  }
  return 12;
}

// gates gs[b, e) on the 2^n amplitudes at p (device)
static void apply_gates_raw(double2* p, uint32_t n, const std::vector<HGate>& gs, size_t b, size_t e) {
  if (b >= e) return;
  std::vector<int> qmap(n);
  std::iota(qmap.begin(), qmap.end(), 0);
  std::vector<LGate> lg;
  lg.reserve(e - b);
  for (size_t i = b; i < e; ++i) lg.push_back(lower_gate(gs[i], qmap));
  Program P;
  plan_program(P, lg, n, default_tile_bits(), 5, false);
  if (P.passes.empty()) return;
  P.upload_all();
  for (size_t i = 0; i < P.passes.size(); ++i) {
    PassDyn dy{};
    dy.raw = reinterpret_cast<double*>(p);
    HIPCHECK(launch_program_pass(P, i, IN_RAW, OUT_RAW, dy, 1, stream()));
  }
  HIPCHECK(hipStreamSynchronize(stream()));
}
static void sv_apply_gates(qsim_sv* sv, const std::vector<HGate>& gs) { apply_gates_raw(sv->ptr(), sv->n, gs, 0, gs.size()); }
}  // namespace qsim

// ====================================================================== CS
namespace qsim {
struct CStage {
  std::vector<uint32_t> gate_ids;
  std::vector<uint32_t> H, O;
  std::vector<uint32_t> b2g;  // as registered (reference layout)
  uint32_t num_groups = 0;
  std::unique_ptr<Program> prog;
};

// Region store (QSIM_STORE_REGIONS): payload words are offsets into one
// virtual reservation: one always-mapped unit, then two halves.  Half h
// holds 2^k regions; region r spans `span` bytes at
// unit + h * half_bytes + r * span and holds the blocks whose
// key qubits read r (key[i] is bit i of r).  Physical units of `unit` bytes
// are mapped at the front of a region as its payload grows and unmapped into
// `pool` once the region has been consumed.
constexpr uint32_t kRegionKeyBits = 4;
constexpr uint32_t kRegionSubs = 64;  // allocation cursors per region
struct RegionStore {
  hipMemAllocationProp prop{};
  hipMemAccessDesc acc{};
  uint64_t unit = 0, half_bytes = 0, va_bytes = 0;
  char* va = nullptr;
  std::vector<hipMemGenericAllocationHandle_t> pool, all;
  struct Half {
    std::vector<uint32_t> key;
    uint64_t span = 0;
    std::vector<std::vector<hipMemGenericAllocationHandle_t>> units;
    std::vector<uint64_t> used;     // words from the region start to the end of its last piece
    std::vector<uint64_t> payload;  // payload words per region (used less the gaps)
  } half[2];
  uint64_t mapped_bytes = 0, peak_mapped_bytes = 0;
  bool zero_mapped = false;  // the unit at offset 0
  // counters of the current stage (QOMPSIM_REGION_LOG=1 prints them)
  uint64_t n_launch = 0, n_retry = 0, n_map = 0, n_unmap = 0;
  double t_map = 0.0, t_unmap = 0.0, t_wait = 0.0;
  RegionStore() = default;
  RegionStore(const RegionStore&) = delete;
  RegionStore& operator=(const RegionStore&) = delete;
  ~RegionStore() {
    if (!va) return;
    (void)hipDeviceSynchronize();
    for (int h = 0; h < 2; ++h)
      for (size_t r = 0; r < half[h].units.size(); ++r)
        for (size_t u = 0; u < half[h].units[r].size(); ++u)
          (void)hipMemUnmap(va + unit + h * half_bytes + r * half[h].span + u * unit, unit);
    if (zero_mapped) (void)hipMemUnmap(va, unit);
    for (auto hd : all) (void)hipMemRelease(hd);
    (void)hipMemAddressFree(va, va_bytes);
  }
};
}  // namespace qsim

struct qsim_cs {
  uint32_t n = 0, B = 0;
  uint64_t nb = 0, cpb = 0, dim = 0;
  qsim_cs_config cfg{};
  std::vector<HGate> gates;
  std::vector<CStage> stages;
  // descriptor tables: the double store's input and output; the region store
  // keeps one (desc[0], encoded in place) and reads a stage batch's input
  // descriptors from a copy in desc_stage
  DBuf desc[2], desc_stage, arena[2], cursors[2], cstats, subbase[2];
  std::vector<uint64_t> h_subbase[2];
  uint64_t cap_words[2] = {0, 0};
  uint32_t nsub[2] = {kSubArenas, kSubArenas};
  DBuf null_flags, nonzero, overflow, bmap, scratch, bytes_tmp;
  uint64_t scratch_amps = 0;
  int cur = 0;
  std::vector<uint8_t> h_null;
  std::vector<uint64_t> bytes_hist;
  uint64_t cur_bytes = 0;
  std::vector<std::array<double, 5>> stats;
  double timers[QSIM_TIMER_COUNT] = {0};
  // events after each launch of a multi-pass stage and the timer each
  // interval counts under (decomp, gate or comp)
  std::vector<hipEvent_t> pev;
  std::vector<int> pkind;
  size_t npev = 0;
  Program io;  // zero-op program over one block (load / readout)
  Partials part;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool regions = false;  // QSIM_STORE_REGIONS
  qsim::RegionStore rs;
  DBuf slot_region;
};

namespace qsim {
// The descriptor table of store half `which` (the region store has one)
static DBuf& cs_desc(qsim_cs* cs, int which) { return cs->regions ? cs->desc[0] : cs->desc[which]; }

// Device bytes the store holds now: descriptor tables, staging buffer and
// arenas (region store: mapped units, the offset-0 unit included).
static uint64_t cs_held_bytes(const qsim_cs* cs) {
  uint64_t b = cs->desc[0].bytes + cs->desc[1].bytes + cs->desc_stage.bytes + cs->scratch.bytes;
  if (cs->regions) b += cs->rs.mapped_bytes + (cs->rs.zero_mapped ? cs->rs.unit : 0);
  else b += cs->arena[0].bytes + cs->arena[1].bytes;
  return b;
}

// Throws when holding `extra` more bytes (and releasing `freed`) would pass
// the configured budget.
static void cs_budget_check(const qsim_cs* cs, uint64_t extra, uint64_t freed, const char* what) {
  if (!cs->cfg.hbm_budget) return;
  const uint64_t after = cs_held_bytes(cs) + extra - std::min(freed, cs_held_bytes(cs));
  if (after > cs->cfg.hbm_budget)
    throw std::runtime_error(std::string("HBM budget of ") + std::to_string(cs->cfg.hbm_budget) +
                             " bytes exceeded (" + what + " would hold " + std::to_string(after) + ")");
}

// Timers of get_total_times (compressed_state.hpp:52-79).  Stage kernels
// count under decomp (DEC->RAW passes), gate (RAW->RAW passes and the fused
// DEC->ENC launch of a single-pass stage, which decodes, applies the gates and
// encodes in one kernel) and comp (RAW->ENC passes), from HIP events on the
// library stream; device allocation, release and host<->device copies of the
// store are wall-clock times.  fused_swap, allgather and hmalloc stay 0 (no
// ranks, no host block buffers).
enum : int { T_DECOMP = 0, T_GATE = 1, T_COMP = 2, T_HMALLOC = 5, T_DMALLOC = 6, T_H2D = 7, T_D2H = 8, T_DFREE = 9 };
struct WallTimer {
  double& acc;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  explicit WallTimer(double& a) : acc(a) {}
  ~WallTimer() { acc += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};
static int pass_timer(int in_mode, int out_mode) {
  return in_mode == IN_DEC && out_mode == OUT_RAW ? T_DECOMP : in_mode == IN_RAW && out_mode == OUT_ENC ? T_COMP : T_GATE;
}
// mark the end of a launch of timer kind k (after cs->ev0 was recorded)
static void cs_mark(qsim_cs* cs, int k) {
  if (cs->npev == cs->pev.size()) {
    hipEvent_t e;
    HIPCHECK(hipEventCreate(&e));
    cs->pev.push_back(e);
    cs->pkind.push_back(0);
  }
  HIPCHECK(hipEventRecord(cs->pev[cs->npev], stream()));
  cs->pkind[cs->npev++] = k;
}
// after the stage's last event: every interval from ev0 through the marks
static void cs_collect_marks(qsim_cs* cs) {
  hipEvent_t prev = cs->ev0;
  for (size_t i = 0; i < cs->npev; ++i) {
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, prev, cs->pev[i]));
    cs->timers[cs->pkind[i]] += ms * 1e-3;
    prev = cs->pev[i];
  }
  cs->npev = 0;
}

static uint64_t words_for_cap(uint64_t bytes) {
  uint64_t w = (bytes + 7) / 8;
  w = (w + kSubArenas - 1) / kSubArenas * kSubArenas;
  return w;
}

// The decoder reads whole dwords around a code field and may touch up to 8
// bytes past a chunk's payload, so every arena keeps 64 bytes of padding.
static void cs_alloc_arena(qsim_cs* cs, int which, uint64_t words) {
  cs_budget_check(cs, (words + 8) * 8, cs->arena[which].bytes, "arena growth");
  {
    WallTimer t(cs->timers[T_DFREE]);
    cs->arena[which].release();
  }
  WallTimer t(cs->timers[T_DMALLOC]);
  cs->arena[which].alloc((words + 8) * 8);
  cs->cap_words[which] = words;
}

// Number of allocation cursors for an encode of `tiles` tiles: about 64
// tiles per cursor keeps the device-scope atomics on one address few.
static uint32_t pick_nsub(uint64_t tiles) {
  uint32_t S = kSubArenas;
  while (S < (uint32_t)kMaxSubArenas && (uint64_t)S * 64 < tiles) S <<= 1;
  return S;
}

// Sub-arena s of output arena `out` spans words [base[s], base[s+1]).
static void cs_upload_subbase(qsim_cs* cs, int out) {
  upload(cs->subbase[out], cs->h_subbase[out].data(), cs->h_subbase[out].size() * sizeof(uint64_t));
}

// Clear the allocation state of output arena `out` for a launch with `nsub`
// cursors.  A first attempt splits the capacity evenly; a retry after an
// overflow keeps the per-sub-arena sizes cs_check_overflow measured.
static void cs_reset_out(qsim_cs* cs, int out, uint32_t nsub, bool keep_partition = false) {
  if (!keep_partition || cs->nsub[out] != nsub || cs->h_subbase[out].size() != (size_t)nsub + 1) {
    cs->nsub[out] = nsub;
    auto& hb = cs->h_subbase[out];
    hb.resize((size_t)nsub + 1);
    const uint64_t per = cs->cap_words[out] / nsub;
    for (uint32_t s = 0; s <= nsub; ++s) hb[s] = (uint64_t)s * per;
    cs_upload_subbase(cs, out);
  }
  HIPCHECK(launch_fill_u64(cs->cursors[out].as<unsigned long long>(), (uint64_t)nsub * 16, 0ull, stream()));
  HIPCHECK(launch_fill_u8(cs->overflow.as<uint8_t>(), sizeof(uint32_t), 0, stream()));
}

static uint64_t cs_read_cursor_words(qsim_cs* cs, int which, uint64_t* maxsub) {
  unsigned long long h[2];
  HIPCHECK(launch_cursor_stats(cs->cursors[which].as<unsigned long long>(), cs->nsub[which],
                               cs->cstats.as<unsigned long long>(), stream()));
  HIPCHECK(hipMemcpyAsync(h, cs->cstats.p, sizeof(h), hipMemcpyDeviceToHost, stream()));
  HIPCHECK(hipStreamSynchronize(stream()));
  if (maxsub) *maxsub = h[1];
  return h[0];
}

// After a launch: false when every tile found room.  On overflow the cursors
// hold each sub-arena's full demand (tiles that did not fit still counted);
// the sub-arenas are resized to that demand plus 1/4 and the arena grows
// only when the total exceeds its capacity, so the retry allocates about what
// the stage needs however unevenly the demand falls across sub-arenas.
static bool cs_check_overflow(qsim_cs* cs, int out) {
  uint32_t ov = 0;
  HIPCHECK(hipMemcpyAsync(&ov, cs->overflow.p, sizeof(ov), hipMemcpyDeviceToHost, stream()));
  HIPCHECK(hipStreamSynchronize(stream()));
  if (!ov) return false;
  const uint32_t ns = cs->nsub[out];
  std::vector<unsigned long long> cur((size_t)ns * 16);
  HIPCHECK(hipMemcpyAsync(cur.data(), cs->cursors[out].p, cur.size() * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, stream()));
  HIPCHECK(hipStreamSynchronize(stream()));
  auto& hb = cs->h_subbase[out];
  hb.assign((size_t)ns + 1, 0);
  uint64_t dmax = 0;
  for (uint32_t s = 0; s < ns; ++s) {
    const uint64_t d = cur[(size_t)s * 16];
    hb[s + 1] = hb[s] + d + d / 4 + 64;
    dmax = std::max(dmax, d + d / 4 + 64);
  }
  // Later stages start from an equal partition again: where device memory
  // allows, grow the arena so that an equal split holds this stage's largest
  // sub-arena demand, otherwise to the measured total only.
  const uint64_t equal_words = dmax * ns;
  if (equal_words > cs->cap_words[out]) {
    size_t free_b = 0, total_b = 0;
    HIPCHECK(hipMemGetInfo(&free_b, &total_b));
    const uint64_t held = cs->cap_words[out] * 8;
    const uint64_t want = words_for_cap(equal_words * 8) * 8;
    const bool in_budget =
        !cs->cfg.hbm_budget || cs_held_bytes(cs) - cs->arena[out].bytes + want + 64 <= cs->cfg.hbm_budget;
    if (want <= held + free_b / 2 && in_budget) cs_alloc_arena(cs, out, want / 8);
    else if (hb[ns] > cs->cap_words[out]) cs_alloc_arena(cs, out, words_for_cap(hb[ns] * 8));
  }
  const uint64_t slack = cs->cap_words[out] - hb[ns];
  hb[ns] += slack;  // leftover capacity goes to the last sub-arena
  cs_upload_subbase(cs, out);
  return true;
}

static PassDyn cs_dyn(qsim_cs* cs, int out) {
  PassDyn dy{};
  dy.null_in = cs->null_flags.as<uint8_t>();
  dy.desc_in = cs_desc(cs, cs->cur).as<uint64_t>();
  dy.desc_out = cs_desc(cs, out).as<uint64_t>();
  if (cs->regions) {
    dy.arena_in = reinterpret_cast<const uint64_t*>(cs->rs.va);
    dy.arena_out = reinterpret_cast<uint64_t*>(cs->rs.va);
  } else {
    dy.arena_in = cs->arena[cs->cur].as<uint64_t>();
    dy.arena_out = cs->arena[out].as<uint64_t>();
  }
  dy.sub_base = cs->subbase[out].as<uint64_t>();
  dy.cursors = cs->cursors[out].as<unsigned long long>();
  dy.sub_mask = cs->regions ? kRegionSubs - 1 : cs->nsub[out] - 1;
  dy.overflow = cs->overflow.as<uint32_t>();
  dy.nonzero = cs->nonzero.as<uint8_t>();
  return dy;
}

// High word of the smallest |x| for which the quantiser's 2*fma(eps_rel,|x|,eps_abs)
// is not finite (the predicate is monotone in the bit pattern of |x|).
static uint32_t codec_overflow_hi(double eps_rel, double eps_abs) {
  auto ovf = [&](uint64_t bits) {
    double x;
    std::memcpy(&x, &bits, sizeof x);
    return !std::isfinite(2.0 * std::fma(eps_rel, x, eps_abs));
  };
  if (ovf(0)) return 0;
  uint64_t lo = 0, hi = 0x7ff0000000000000ull;
  while (hi - lo > 1) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (ovf(mid)) hi = mid;
    else lo = mid;
  }
  return (uint32_t)(hi >> 32);
}

static void cs_fill_codec_fields(qsim_cs* cs, Program& P, uint32_t khi, int zero_mode) {
  const uint32_t ovf_hi = codec_overflow_hi(cs->cfg.eps_rel, cs->cfg.eps_abs);
  for (auto& a : P.passes) {
    a.ovf_hi = ovf_hi;
    a.B = cs->B;
    a.khi = khi;
    a.cpb = cs->cpb;
    a.eps_rel = cs->cfg.eps_rel;
    a.eps_abs = cs->cfg.eps_abs;
    a.guard = cs->cfg.guard_bits;
    a.zero_mode = zero_mode;
    a.zero_thr = cs->cfg.zero_threshold;
    if (cs->cfg.eps_abs > 0.0) {
      a.e_abs = std::ilogb(2.0 * cs->cfg.eps_abs);
      a.p_abs = std::ldexp(1.0, a.e_abs);
    } else {
      a.e_abs = INT32_MIN;
      a.p_abs = 0.0;
    }
  }
}

static void cs_commit(qsim_cs* cs, int out) {
  HIPCHECK(launch_commit_null(cs->null_flags.as<uint8_t>(), cs->nonzero.as<uint8_t>(), cs->nb, stream()));
  HIPCHECK(hipMemcpyAsync(cs->h_null.data(), cs->null_flags.p, cs->nb, hipMemcpyDeviceToHost, stream()));
  HIPCHECK(hipStreamSynchronize(stream()));
  cs->cur = out;
}

static uint64_t rs_current_bytes(qsim_cs* cs);
static uint64_t cs_current_bytes(qsim_cs* cs) {
  if (cs->regions) return rs_current_bytes(cs);
  const uint64_t words = cs_read_cursor_words(cs, cs->cur, nullptr);
  uint64_t nonnull = 0;
  for (auto v : cs->h_null) nonnull += v ? 0 : 1;
  return 8 * (words + nonnull * cs->cpb);
}

// Encode the blocks held in scratch (block i of the batch at scratch +
// i*2^B) into the output arena; block ids from h_bmap.
static void cs_encode_scratch(qsim_cs* cs, int out, const std::vector<uint32_t>& ids) {
  upload(cs->bmap, ids.data(), ids.size() * sizeof(uint32_t));
  PassDyn dy = cs_dyn(cs, out);
  dy.raw = cs->scratch.as<double>();
  dy.bmap = cs->bmap.as<uint32_t>();
  HIPCHECK(launch_program_pass(cs->io, 0, IN_RAW, OUT_ENC, dy, ids.size(), stream()));
}

static void cs_decode_to_scratch(qsim_cs* cs, const std::vector<uint32_t>& ids) {
  upload(cs->bmap, ids.data(), ids.size() * sizeof(uint32_t));
  PassDyn dy = cs_dyn(cs, 1 - cs->cur);
  dy.raw = cs->scratch.as<double>();
  dy.bmap = cs->bmap.as<uint32_t>();
  HIPCHECK(launch_program_pass(cs->io, 0, IN_DEC, OUT_RAW, dy, ids.size(), stream()));
}

static uint64_t cs_blocks_per_batch(qsim_cs* cs) {
  return std::max<uint64_t>(1, std::min<uint64_t>(cs->nb, cs->scratch_amps >> cs->B));
}

static void cs_load_regions(qsim_cs* cs, const double* host, const double2* dev, uint64_t basis);

// Replace the whole state by a host vector (nullptr: basis state `basis`).
static void cs_load(qsim_cs* cs, const double* host, uint64_t basis) {
  if (cs->regions) return cs_load_regions(cs, host, nullptr, basis);
  const int out = 1 - cs->cur;
  for (int attempt = 0;; ++attempt) {
    HIPCHECK(launch_fill_u8(cs->null_flags.as<uint8_t>(), cs->nb, 1, stream()));
    HIPCHECK(launch_fill_u8(cs->nonzero.as<uint8_t>(), cs->nb, 0, stream()));
    {
      const PassArgs& io = cs->io.passes[0];
      const uint32_t tb = io.t;
      cs_reset_out(cs, out, pick_nsub((host ? cs->nb : 1ull) << (cs->B - tb)), attempt > 0);
    }
    if (host) {
      const uint64_t bpb = cs_blocks_per_batch(cs);
      for (uint64_t b0 = 0; b0 < cs->nb; b0 += bpb) {
        const uint64_t cnt = std::min(bpb, cs->nb - b0);
        {
          WallTimer t(cs->timers[T_H2D]);
          HIPCHECK(hipMemcpyAsync(cs->scratch.p, host + 2 * (b0 << cs->B), (cnt << cs->B) * 16, hipMemcpyHostToDevice,
                                  stream()));
          HIPCHECK(hipStreamSynchronize(stream()));
        }
        std::vector<uint32_t> ids(cnt);
        std::iota(ids.begin(), ids.end(), (uint32_t)b0);
        cs_encode_scratch(cs, out, ids);
        HIPCHECK(hipStreamSynchronize(stream()));
      }
    } else {
      const uint32_t b = (uint32_t)(basis >> cs->B);
      HIPCHECK(launch_fill_basis(cs->scratch.as<double2>(), 1ull << cs->B, basis & ((1ull << cs->B) - 1), stream()));
      cs_encode_scratch(cs, out, {b});
    }
    if (!cs_check_overflow(cs, out)) break;
    if (attempt > 8) throw std::runtime_error("compressed store: arena growth did not converge");
  }
  cs_commit(cs, out);
  cs->cur_bytes = cs_current_bytes(cs);
}

// Replace the whole state by a device vector of 2^n amplitudes: blocks are
// encoded straight from it (block b = amplitudes [b 2^B, (b+1) 2^B)).
static void cs_load_device(qsim_cs* cs, const double2* dev) {
  if (cs->regions) return cs_load_regions(cs, nullptr, dev, 0);
  const int out = 1 - cs->cur;
  std::vector<uint32_t> ids(cs->nb);
  std::iota(ids.begin(), ids.end(), 0u);
  for (int attempt = 0;; ++attempt) {
    HIPCHECK(launch_fill_u8(cs->null_flags.as<uint8_t>(), cs->nb, 1, stream()));
    HIPCHECK(launch_fill_u8(cs->nonzero.as<uint8_t>(), cs->nb, 0, stream()));
    {
      const PassArgs& io = cs->io.passes[0];
      const uint32_t tb = io.t;
      cs_reset_out(cs, out, pick_nsub(cs->nb << (cs->B - tb)), attempt > 0);
    }
    upload(cs->bmap, ids.data(), ids.size() * sizeof(uint32_t));
    PassDyn dy = cs_dyn(cs, out);
    dy.raw = (double*)dev;
    dy.bmap = cs->bmap.as<uint32_t>();
    HIPCHECK(launch_program_pass(cs->io, 0, IN_RAW, OUT_ENC, dy, ids.size(), stream()));
    if (!cs_check_overflow(cs, out)) break;
    if (attempt > 8) throw std::runtime_error("compressed store: arena growth did not converge");
  }
  cs_commit(cs, out);
  cs->cur_bytes = cs_current_bytes(cs);
}

// Decode the whole state into a device vector of 2^n amplitudes.
static void cs_read_device(qsim_cs* cs, double2* dev) {
  std::vector<uint32_t> ids(cs->nb);
  std::iota(ids.begin(), ids.end(), 0u);
  upload(cs->bmap, ids.data(), ids.size() * sizeof(uint32_t));
  PassDyn dy = cs_dyn(cs, 1 - cs->cur);
  dy.raw = (double*)dev;
  dy.bmap = cs->bmap.as<uint32_t>();
  HIPCHECK(launch_program_pass(cs->io, 0, IN_DEC, OUT_RAW, dy, ids.size(), stream()));
}


// make_block_groups (compressed_state.hpp:585-671) for one stage.
static void stage_groups(qsim_cs* cs, CStage& S) {
  std::set<uint32_t> inner;
  for (auto gid : S.gate_ids) {
    for (auto q : cs->gates[gid].t)
      if (q >= cs->B) inner.insert(q);
    for (auto q : cs->gates[gid].c)
      if (q >= cs->B) inner.insert(q);
  }
  S.H.assign(inner.begin(), inner.end());
  S.O.clear();
  for (uint32_t q = cs->B; q < cs->n; ++q)
    if (!inner.count(q)) S.O.push_back(q);
}

static std::vector<uint32_t> canonical_b2g(qsim_cs* cs, const CStage& S) {
  std::vector<uint32_t> b2g(cs->nb);
  for (uint64_t b = 0; b < cs->nb; ++b) {
    uint32_t g = 0;
    for (size_t i = 0; i < S.O.size(); ++i) g |= (uint32_t)((b >> (S.O[i] - cs->B)) & 1u) << i;
    b2g[b] = g;
  }
  return b2g;
}

static void build_stage_program(qsim_cs* cs, CStage& S) {
  const uint32_t k = (uint32_t)S.H.size();
  const uint32_t L = cs->B + k;
  if (L > (uint32_t)kMaxLocalBits) throw std::invalid_argument("stage touches too many high qubits");
  std::vector<int> qmap(cs->n, -1);
  for (uint32_t q = 0; q < cs->B; ++q) qmap[q] = (int)q;
  for (uint32_t i = 0; i < k; ++i) qmap[S.H[i]] = (int)(cs->B + i);
  std::vector<LGate> lg;
  for (auto gid : S.gate_ids) lg.push_back(lower_gate(cs->gates[gid], qmap));
  S.prog.reset(new Program());
  const uint32_t tb = cs->cfg.tile_bits ? cs->cfg.tile_bits : default_tile_bits();
  plan_program(*S.prog, lg, L, std::min<uint32_t>(tb, L), kChunkAmpsLog2, true, true);
  cs_fill_codec_fields(cs, *S.prog, k, cs->cfg.zero_mode);
  S.prog->upload_all();
  const uint64_t need = (1ull << L);
  if (S.prog->passes.size() > 1 && cs->scratch_amps < need) {
    cs_budget_check(cs, need * 16, cs->scratch.bytes, "staging buffer growth");
    WallTimer t(cs->timers[T_DMALLOC]);
    cs->scratch.alloc(need * 16);
    cs->scratch_amps = need;
  }
}

// partition_circuit (compressed_state.hpp:509-583)
static std::vector<std::vector<uint32_t>> partition(qsim_cs* cs, uint32_t inner_size) {
  std::vector<std::vector<uint32_t>> stages;
  std::vector<uint32_t> current, gidx;
  const uint32_t thr = std::max<uint32_t>(inner_size, 2u);
  for (uint32_t i = 0; i < cs->gates.size(); ++i) {
    const HGate& g = cs->gates[i];
    std::vector<uint32_t> gt;
    for (auto q : g.t)
      if (q >= cs->B) gt.push_back(q);
    for (auto q : g.c)
      if (q >= cs->B) gt.push_back(q);
    std::vector<uint32_t> tmp = gidx;
    for (auto q : gt)
      if (std::find(tmp.begin(), tmp.end(), q) == tmp.end()) tmp.push_back(q);
    if (tmp.size() > thr) {
      stages.push_back(current);
      current.clear();
      tmp = gt;
    }
    gidx = tmp;
    current.push_back(i);
  }
  if (!current.empty()) stages.push_back(current);
  return stages;
}

static void cs_run_stage_regions(qsim_cs* cs, uint32_t si);
static void cs_run_stage(qsim_cs* cs, uint32_t si) {
  if (cs->regions) return cs_run_stage_regions(cs, si);
  CStage& S = cs->stages[si];
  Program& P = *S.prog;
  const uint32_t k = (uint32_t)S.H.size();
  const uint32_t L = cs->B + k;
  const uint64_t G = 1ull << k;
  // build_stage_group_execution_plan (compressed_state.hpp:797-837) on the
  // canonical groups: a group is active when it holds a non-null block.
  const auto b2g = canonical_b2g(cs, S);
  const uint64_t ng = 1ull << S.O.size();
  std::vector<uint8_t> active(ng, 0);
  for (uint64_t b = 0; b < cs->nb; ++b)
    if (!cs->h_null[b]) active[b2g[b]] = 1;
  std::vector<uint32_t> bm;
  uint64_t nact = 0;
  for (uint64_t g = 0; g < ng; ++g) {
    if (!active[g]) continue;
    uint64_t obits = 0;
    for (size_t i = 0; i < S.O.size(); ++i) obits |= ((g >> i) & 1ull) << (S.O[i] - cs->B);
    for (uint64_t j = 0; j < G; ++j) {
      uint64_t b = obits;
      for (uint32_t i = 0; i < k; ++i) b |= ((j >> i) & 1ull) << (S.H[i] - cs->B);
      bm.push_back((uint32_t)b);
    }
    ++nact;
  }
  const uint64_t bytes_in = cs->cur_bytes;
  if (nact == 0 || P.passes.empty()) {
    cs->bytes_hist.push_back(bytes_in);
    cs->stats.push_back({0.0, (double)bytes_in, (double)bytes_in, 0.0, 0.0});
    return;
  }
  upload(cs->bmap, bm.data(), bm.size() * sizeof(uint32_t));
  const int out = 1 - cs->cur;
  float ms = 0.f;
  // A single wave pass (the persistent DEC->ENC kernel) on the double store
  // gives each wave of its grid a sub-arena of its own.
  const uint64_t last_tiles = nact << (L - P.passes.back().t);
  const bool own = P.passes.size() == 1 && P.passes[0].engine == ENGINE_WAVE && !cs->regions &&
                   wave_persistent_waves(last_tiles) <= (uint64_t)kMaxSubArenas;
  const uint32_t nsub = own ? (uint32_t)wave_persistent_waves(last_tiles) : pick_nsub(last_tiles);
  for (int attempt = 0;; ++attempt) {
    HIPCHECK(launch_fill_u8(cs->nonzero.as<uint8_t>(), cs->nb, 0, stream()));
    cs_reset_out(cs, out, nsub, attempt > 0);
    HIPCHECK(hipEventRecord(cs->ev0, stream()));
    cs->npev = 0;
    PassDyn dy = cs_dyn(cs, out);
    dy.wave_subs = own ? 1u : 0u;
    if (P.passes.size() == 1) {
      dy.bmap = cs->bmap.as<uint32_t>();
      HIPCHECK(launch_program_pass(P, 0, IN_DEC, OUT_ENC, dy, nact, stream()));
      cs_mark(cs, T_GATE);
    } else {
      const uint64_t gpb = std::max<uint64_t>(1, cs->scratch_amps >> L);
      dy.raw = cs->scratch.as<double>();
      for (uint64_t a0 = 0; a0 < nact; a0 += gpb) {
        const uint64_t cnt = std::min(gpb, nact - a0);
        dy.bmap = cs->bmap.as<uint32_t>() + a0 * G;
        for (size_t p = 0; p < P.passes.size(); ++p) {
          const int in = p == 0 ? IN_DEC : IN_RAW;
          const int om = p + 1 == P.passes.size() ? OUT_ENC : OUT_RAW;
          HIPCHECK(launch_program_pass(P, p, in, om, dy, cnt, stream()));
          cs_mark(cs, pass_timer(in, om));
        }
      }
    }
    HIPCHECK(hipEventRecord(cs->ev1, stream()));
    if (!cs_check_overflow(cs, out)) break;
    if (attempt > 8) throw std::runtime_error("compressed store: arena growth did not converge");
  }
  HIPCHECK(hipEventElapsedTime(&ms, cs->ev0, cs->ev1));
  cs_collect_marks(cs);
  cs_commit(cs, out);
  const uint64_t bytes_out = cs_current_bytes(cs);
  cs->cur_bytes = bytes_out;
  cs->bytes_hist.push_back(bytes_out);
  cs->stats.push_back({ms * 1e-3, (double)bytes_in, (double)bytes_out, (double)(nact << L), (double)P.passes.size()});
}

// ------------------------------------------------------------ region store
static void rs_init(qsim_cs* cs) {
  RegionStore& R = cs->rs;
  int dev = 0, vmm = 0;
  HIPCHECK(hipGetDevice(&dev));
  HIPCHECK(hipDeviceGetAttribute(&vmm, hipDeviceAttributeVirtualMemoryManagementSupported, dev));
  if (!vmm) throw std::runtime_error("region store: the device does not support virtual memory management");
  R.prop.type = hipMemAllocationTypePinned;
  R.prop.location.type = hipMemLocationTypeDevice;
  R.prop.location.id = dev;
  R.acc.location = R.prop.location;
  R.acc.flags = hipMemAccessFlagsProtReadWrite;
  size_t gran = 0;
  HIPCHECK(hipMemGetAllocationGranularity(&gran, &R.prop, hipMemAllocationGranularityMinimum));
  if (gran == 0) gran = 4096;
  const uint64_t raw = cs->dim * 16;
  // about 1024 units per raw state (256 MiB from 34 qubits up, 64 KiB at 22):
  // unmapping costs about 0.25 ms a unit whatever its size
  uint64_t unit = 64ull << 10;
  while (unit < (256ull << 20) && unit < raw / 1024) unit <<= 1;
  R.unit = (unit + gran - 1) / gran * gran;
  // a half holds 2^k regions of at most 9/8 raw / 2^k + 2 units each
  R.half_bytes = (raw + raw / 8 + (2ull << kRegionKeyBits) * R.unit + R.unit - 1) / R.unit * R.unit;
  R.va_bytes = R.unit + 2 * R.half_bytes;
  if (R.va_bytes / 8 >= (1ull << 40)) throw std::invalid_argument("region store: state too large");
  HIPCHECK(hipMemAddressReserve((void**)&R.va, R.va_bytes, R.unit, nullptr, 0));
  // The first unit stays mapped: descriptors of chunks without payload hold
  // offset 0, and the decoders read a few words there before masking them.
  hipMemGenericAllocationHandle_t hd;
  HIPCHECK(hipMemCreate(&hd, R.unit, &R.prop, 0));
  R.all.push_back(hd);
  HIPCHECK(hipMemMap(R.va, R.unit, 0, hd, 0));
  HIPCHECK(hipMemSetAccess(R.va, R.unit, &R.acc, 1));
  R.zero_mapped = true;
  HIPCHECK(hipMemsetAsync(R.va, 0, R.unit, stream()));
  HIPCHECK(hipStreamSynchronize(stream()));
}

static char* rs_region_ptr(qsim_cs* cs, int h, uint32_t r) {
  return cs->rs.va + cs->rs.unit + (uint64_t)h * cs->rs.half_bytes + (uint64_t)r * cs->rs.half[h].span;
}

// Unmap region r of half h into the pool.  With `wait`, waits for the
// stream first (queued work may still read the region); without, the caller
// knows that no queued work reads it.
static void rs_release_region(qsim_cs* cs, int h, uint32_t r, bool wait = true) {
  RegionStore& R = cs->rs;
  auto& u = R.half[h].units[r];
  if (!u.empty()) {
    if (wait) HIPCHECK(hipStreamSynchronize(stream()));
    const auto t0 = std::chrono::steady_clock::now();
    char* base = rs_region_ptr(cs, h, r);
    for (size_t i = 0; i < u.size(); ++i) {
      HIPCHECK(hipMemUnmap(base + i * R.unit, R.unit));
      R.pool.push_back(u[i]);
    }
    R.mapped_bytes -= u.size() * R.unit;
    R.n_unmap += u.size();
    u.clear();
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    R.t_unmap += dt;
    cs->timers[T_DFREE] += dt;
  }
  R.half[h].used[r] = 0;
  R.half[h].payload[r] = 0;
}

static void rs_release_half(qsim_cs* cs, int h) {
  for (uint32_t r = 0; r < cs->rs.half[h].units.size(); ++r) rs_release_region(cs, h, r);
}

// Empty half h and lay it out as 2^|key| regions keyed by `key`.
static void rs_set_layout(qsim_cs* cs, int h, const std::vector<uint32_t>& key) {
  RegionStore& R = cs->rs;
  rs_release_half(cs, h);
  auto& H = R.half[h];
  H.key = key;
  // An encode lays out at most the raw size of its blocks (capped first
  // estimate) or exactly their demand (retry), so a region's pieces never
  // pass its raw size; the span adds 1/8 and a unit of margin.
  const uint64_t rr = (cs->dim * 16) >> key.size();
  H.span = (rr + rr / 8 + 2 * R.unit - 1) / R.unit * R.unit;
  H.units.assign(1u << key.size(), {});
  H.used.assign(1u << key.size(), 0);
  H.payload.assign(1u << key.size(), 0);
}

// Map units at the front of region r until it holds `words` payload words
// plus the decoder's 64 bytes of padding.
static void rs_map(qsim_cs* cs, int h, uint32_t r, uint64_t words) {
  RegionStore& R = cs->rs;
  auto& H = R.half[h];
  const uint64_t need = (words + 8) * 8;
  if (need > H.span) throw std::logic_error("region store: pieces past the region's span");
  auto& u = H.units[r];
  char* base = rs_region_ptr(cs, h, r);
  if (u.size() * R.unit >= need) return;
  const auto t0 = std::chrono::steady_clock::now();
  while (u.size() * R.unit < need) {
    hipMemGenericAllocationHandle_t hd;
    cs_budget_check(cs, R.unit, 0, "region mapping");
    if (!R.pool.empty()) {
      hd = R.pool.back();
      R.pool.pop_back();
    } else {
      const hipError_t e = hipMemCreate(&hd, R.unit, &R.prop, 0);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        throw std::runtime_error(std::string("region store: out of device memory (") + hipGetErrorString(e) + ")");
      }
      R.all.push_back(hd);
    }
    char* p = base + u.size() * R.unit;
    HIPCHECK(hipMemMap(p, R.unit, 0, hd, 0));
    HIPCHECK(hipMemSetAccess(p, R.unit, &R.acc, 1));
    u.push_back(hd);
    R.mapped_bytes += R.unit;
    R.peak_mapped_bytes = std::max(R.peak_mapped_bytes, R.mapped_bytes);
    ++R.n_map;
  }
  const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  R.t_map += dt;
  cs->timers[T_DMALLOC] += dt;
}

static uint32_t key_value(const qsim_cs* cs, uint64_t block, const std::vector<uint32_t>& key) {
  uint32_t v = 0;
  for (size_t i = 0; i < key.size(); ++i) v |= (uint32_t)((block >> (key[i] - cs->B)) & 1u) << i;
  return v;
}

// Up to kRegionKeyBits qubits >= B outside `excl`, the highest first.
static std::vector<uint32_t> rs_pick_key(const qsim_cs* cs, const std::set<uint32_t>& excl) {
  std::vector<uint32_t> key;
  for (uint32_t q = cs->n; q-- > cs->B && key.size() < kRegionKeyBits;)
    if (!excl.count(q)) key.push_back(q);
  std::sort(key.begin(), key.end());
  return key;
}

// One encode into half h.  Region r's next est[r] words are cut into
// kRegionSubs pieces, one allocation cursor each (sub-arena r kRegionSubs + i
// of the launch; a tile t takes i = t mod kRegionSubs), so the waves working
// on one group spread their atomics.  Pieces are laid out in order; what a
// piece leaves unused stays a gap.  On overflow the cursors hold every
// piece's demand (tiles that did not fit included) and the retry lays the
// pieces out at exactly that demand.  `during` runs once, while the first
// attempt's kernels are in flight.
template <class F, class D>
static void rs_encode(qsim_cs* cs, int h, const std::vector<uint64_t>& est, F&& launch, D&& during,
                      std::vector<uint64_t>* demand) {
  auto& H = cs->rs.half[h];
  const uint32_t K = (uint32_t)H.used.size(), M = kRegionSubs, S = K * M;
  std::vector<uint64_t> piece((size_t)S), w0(K);
  for (uint32_t r = 0; r < K; ++r) {
    w0[r] = (uint64_t)(rs_region_ptr(cs, h, r) - cs->rs.va) / 8;
    for (uint32_t i = 0; i < M; ++i) piece[(size_t)r * M + i] = (est[r] + M - 1) / M;
  }
  std::vector<unsigned long long> cur((size_t)2 * S * 16);
  std::vector<uint64_t> sb((size_t)2 * S + 1, 0);
  bool ran = false;
  for (int attempt = 0;; ++attempt) {
    for (uint32_t r = 0; r < K; ++r) {
      uint64_t w = w0[r] + H.used[r];
      for (uint32_t i = 0; i < M; ++i) {
        const size_t s = (size_t)r * M + i;
        sb[2 * s] = w;
        w += piece[s];
        sb[2 * s + 1] = w;
      }
      rs_map(cs, h, r, w - w0[r]);
    }
    upload(cs->subbase[h], sb.data(), sb.size() * sizeof(uint64_t));
    HIPCHECK(launch_fill_u64(cs->cursors[h].as<unsigned long long>(), (uint64_t)2 * S * 16, 0ull, stream()));
    HIPCHECK(launch_fill_u8(cs->overflow.as<uint8_t>(), sizeof(uint32_t), 0, stream()));
    launch();
    ++cs->rs.n_launch;
    if (!ran) {
      ran = true;
      during();
    }
    uint32_t ov = 0;
    HIPCHECK(hipMemcpyAsync(&ov, cs->overflow.p, sizeof(ov), hipMemcpyDeviceToHost, stream()));
    HIPCHECK(hipMemcpyAsync(cur.data(), cs->cursors[h].p, cur.size() * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, stream()));
    const auto t0 = std::chrono::steady_clock::now();
    HIPCHECK(hipStreamSynchronize(stream()));
    cs->rs.t_wait += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!ov) {
      for (uint32_t r = 0; r < K; ++r) {
        uint64_t tot = 0, end = H.used[r];
        for (uint32_t i = 0; i < M; ++i) {
          const size_t s = (size_t)r * M + i;
          const uint64_t d = cur[2 * s * 16];
          tot += d;
          if (d) end = sb[2 * s] + d - w0[r];
        }
        H.used[r] = end;
        H.payload[r] += tot;
        if (demand) (*demand)[r] = tot;
      }
      return;
    }
    if (attempt > 8) throw std::runtime_error("region store: allocation did not converge");
    ++cs->rs.n_retry;
    // the kernels are deterministic: a retry asks for the same sizes
    for (size_t s = 0; s < S; ++s) piece[s] = cur[2 * s * 16];
  }
}

static uint64_t rs_current_bytes(qsim_cs* cs) {
  uint64_t words = 0, nonnull = 0;
  for (auto u : cs->rs.half[cs->cur].payload) words += u;
  for (auto v : cs->h_null) nonnull += v ? 0 : 1;
  return 8 * (words + nonnull * cs->cpb);
}

// Load for the region store: the old state is dropped first, the new one is
// laid out for the first registered stage.  host / dev: the full vector;
// otherwise the basis state `basis`.
static void cs_load_regions(qsim_cs* cs, const double* host, const double2* dev, uint64_t basis) {
  const int out = 1 - cs->cur;
  rs_release_half(cs, cs->cur);
  std::set<uint32_t> excl;
  if (!cs->stages.empty()) excl.insert(cs->stages[0].H.begin(), cs->stages[0].H.end());
  const std::vector<uint32_t> key = rs_pick_key(cs, excl);
  rs_set_layout(cs, out, key);
  const uint32_t K = 1u << key.size();
  HIPCHECK(launch_fill_u8(cs->null_flags.as<uint8_t>(), cs->nb, 1, stream()));
  HIPCHECK(launch_fill_u8(cs->nonzero.as<uint8_t>(), cs->nb, 0, stream()));
  const bool whole = host || dev;
  std::vector<uint32_t> ids;
  if (whole) {
    ids.resize(cs->nb);
    std::iota(ids.begin(), ids.end(), 0u);
  } else {
    ids.push_back((uint32_t)(basis >> cs->B));
  }
  std::vector<uint32_t> sr(ids.size());
  for (size_t i = 0; i < ids.size(); ++i) sr[i] = key_value(cs, ids[i], key);
  upload(cs->slot_region, sr.data(), sr.size() * sizeof(uint32_t));
  upload(cs->bmap, ids.data(), ids.size() * sizeof(uint32_t));
  // first guess: a quarter of the raw size of a full load (at most a quarter
  // of the free device memory in all), a little for a basis state; a retry
  // maps the measured demand
  uint64_t guess = 64;
  if (whole) {
    size_t free_b = 0, total_b = 0;
    HIPCHECK(hipMemGetInfo(&free_b, &total_b));
    guess = std::min<uint64_t>(((cs->dim * 2) >> key.size()) / 4, (uint64_t)free_b / 4 / 8 / K);
  }
  std::vector<uint64_t> est(K, guess);
  rs_encode(cs, out, est, [&] {
    PassDyn dy = cs_dyn(cs, out);
    dy.slot_region = cs->slot_region.as<uint32_t>();
    dy.bmap = cs->bmap.as<uint32_t>();
    if (dev) {
      dy.raw = (double*)dev;
      HIPCHECK(launch_program_pass(cs->io, 0, IN_RAW, OUT_ENC, dy, ids.size(), stream()));
    } else if (host) {
      dy.raw = cs->scratch.as<double>();
      const uint64_t bpb = cs_blocks_per_batch(cs);
      for (uint64_t b0 = 0; b0 < cs->nb; b0 += bpb) {
        const uint64_t cnt = std::min(bpb, cs->nb - b0);
        {
          WallTimer t(cs->timers[T_H2D]);
          HIPCHECK(hipMemcpyAsync(cs->scratch.p, host + 2 * (b0 << cs->B), (cnt << cs->B) * 16, hipMemcpyHostToDevice,
                                  stream()));
          HIPCHECK(hipStreamSynchronize(stream()));
        }
        dy.bmap = cs->bmap.as<uint32_t>() + b0;
        dy.slot_region = cs->slot_region.as<uint32_t>() + b0;
        HIPCHECK(launch_program_pass(cs->io, 0, IN_RAW, OUT_ENC, dy, cnt, stream()));
        HIPCHECK(hipStreamSynchronize(stream()));
      }
    } else {
      dy.raw = cs->scratch.as<double>();
      HIPCHECK(launch_fill_basis(cs->scratch.as<double2>(), 1ull << cs->B, basis & ((1ull << cs->B) - 1), stream()));
      HIPCHECK(launch_program_pass(cs->io, 0, IN_RAW, OUT_ENC, dy, 1, stream()));
    }
  }, [] {}, nullptr);
  cs_commit(cs, out);
  cs->cur_bytes = rs_current_bytes(cs);
}

// One stage on the region store.  Active groups run in batches keyed by
// qubits the stage does not touch, ordered so that the input's key qubits
// are the most significant: after batch j, every input region whose last
// live block belonged to a batch <= j is unmapped.  The output is routed to
// regions keyed by qubits neither this stage nor the next one touches (the
// first stage follows the last), so the next stage consumes it region by
// region.  Peak: the state, plus the regions in flight, plus what the
// running batch writes.
static void cs_run_stage_regions(qsim_cs* cs, uint32_t si) {
  CStage& S = cs->stages[si];
  Program& P = *S.prog;
  RegionStore& R = cs->rs;
  const uint32_t k = (uint32_t)S.H.size();
  const uint32_t L = cs->B + k;
  const uint64_t G = 1ull << k;
  const int in = cs->cur, out = 1 - cs->cur;
  const uint64_t bytes_in = cs->cur_bytes;
  const std::set<uint32_t> T(S.H.begin(), S.H.end());
  // consumer key: free qubits in the low bits, the input key's untouched
  // qubits (in their order) in the high bits
  const auto& kin = R.half[in].key;
  std::vector<uint32_t> compat, ckey;
  for (auto q : kin)
    if (!T.count(q)) compat.push_back(q);
  for (uint32_t q = cs->n; q-- > cs->B && ckey.size() + compat.size() < kRegionKeyBits;)
    if (!T.count(q) && std::find(compat.begin(), compat.end(), q) == compat.end()) ckey.push_back(q);
  std::reverse(ckey.begin(), ckey.end());
  ckey.insert(ckey.end(), compat.begin(), compat.end());
  const CStage& N = cs->stages[si + 1 < cs->stages.size() ? si + 1 : 0];
  std::set<uint32_t> excl = T;
  excl.insert(N.H.begin(), N.H.end());
  const std::vector<uint32_t> okey = rs_pick_key(cs, excl);
  const uint32_t Kout = 1u << okey.size(), nbatch = 1u << ckey.size();

  // per-block payload words of the input (descriptor widths)
  std::vector<unsigned long long> bb(cs->nb);
  cs->bytes_tmp.ensure(cs->nb * 8);
  HIPCHECK(launch_block_bytes(cs_desc(cs, in).as<uint64_t>(), cs->null_flags.as<uint8_t>(), cs->nb, cs->cpb,
                              cs->bytes_tmp.as<unsigned long long>(), stream()));
  HIPCHECK(hipMemcpyAsync(bb.data(), cs->bytes_tmp.p, cs->nb * 8, hipMemcpyDeviceToHost, stream()));
  HIPCHECK(hipStreamSynchronize(stream()));

  // active groups (build_stage_group_execution_plan, compressed_state.hpp:797-837) by batch
  const uint64_t ng = 1ull << S.O.size();
  std::vector<std::vector<uint64_t>> bgroups(nbatch);
  {
    std::vector<uint8_t> active(ng, 0);
    const auto b2g = canonical_b2g(cs, S);
    for (uint64_t b = 0; b < cs->nb; ++b)
      if (!cs->h_null[b]) active[b2g[b]] = 1;
    for (uint64_t g = 0; g < ng; ++g) {
      if (!active[g]) continue;
      uint64_t obits = 0;
      for (size_t i = 0; i < S.O.size(); ++i) obits |= ((g >> i) & 1ull) << (S.O[i] - cs->B);
      bgroups[key_value(cs, obits, ckey)].push_back(obits);
    }
  }
  std::vector<int64_t> last_use(R.half[in].used.size(), -1);
  std::vector<uint32_t> bm, sr;
  std::vector<uint64_t> slot0(nbatch + 1, 0);
  std::vector<uint64_t> est_in((size_t)nbatch * Kout, 0), raw_out((size_t)nbatch * Kout, 0);
  for (uint32_t j = 0; j < nbatch; ++j) {
    for (uint64_t obits : bgroups[j]) {
      const uint32_t ro = key_value(cs, obits, okey);
      raw_out[(size_t)j * Kout + ro] += G * cs->cpb * 64;
      for (uint64_t jj = 0; jj < G; ++jj) {
        uint64_t b = obits;
        for (uint32_t i = 0; i < k; ++i) b |= ((jj >> i) & 1ull) << (S.H[i] - cs->B);
        bm.push_back((uint32_t)b);
        if (!cs->h_null[b]) {
          last_use[key_value(cs, b, kin)] = j;
          est_in[(size_t)j * Kout + ro] += bb[b] / 8 - cs->cpb;
        }
      }
      sr.push_back(ro);
    }
    slot0[j + 1] = sr.size();
  }
  const uint64_t nact = sr.size();
  if (nact == 0 || P.passes.empty()) {
    cs->bytes_hist.push_back(bytes_in);
    cs->stats.push_back({0.0, (double)bytes_in, (double)bytes_in, 0.0, 0.0});
    return;
  }
  upload(cs->bmap, bm.data(), bm.size() * sizeof(uint32_t));
  upload(cs->slot_region, sr.data(), sr.size() * sizeof(uint32_t));
  HIPCHECK(launch_fill_u8(cs->nonzero.as<uint8_t>(), cs->nb, 0, stream()));
  rs_set_layout(cs, out, okey);
  for (uint32_t r = 0; r < last_use.size(); ++r)
    if (last_use[r] < 0) rs_release_region(cs, in, r);

  R.n_launch = R.n_retry = R.n_map = R.n_unmap = 0;
  R.t_map = R.t_unmap = R.t_wait = 0.0;
  const auto tw0 = std::chrono::steady_clock::now();
  HIPCHECK(hipEventRecord(cs->ev0, stream()));
  cs->npev = 0;
  double ratio = 1.25;
  std::vector<uint64_t> est(Kout), dem(Kout);
  // input regions consumed by the batches so far: unmapped while the next
  // batch runs (no later batch reads them)
  std::vector<uint32_t> pending;
  auto release_pending = [&] {
    for (auto r : pending) rs_release_region(cs, in, r, false);
    pending.clear();
  };
  for (uint32_t j = 0; j < nbatch; ++j) {
    const uint64_t a0 = slot0[j], a1 = slot0[j + 1];
    if (a1 == a0) release_pending();
    if (a1 > a0) {
      uint64_t ein = 0;
      // QOMPSIM_REGION_TIGHT=1 starts every batch from empty pieces, so that
      // each encode retries at the measured demand (for tests)
      const char* tight = std::getenv("QOMPSIM_REGION_TIGHT");
      for (uint32_t r = 0; r < Kout; ++r) {
        const uint64_t e = est_in[(size_t)j * Kout + r];
        ein += e;
        est[r] = std::min<uint64_t>((uint64_t)((double)e * ratio) + 64, raw_out[(size_t)j * Kout + r]);
        if (tight && tight[0] == '1') est[r] = 0;
      }
      // the batch's input descriptors, copied out before its output is
      // encoded over them (a retry reads the copy again)
      cs->desc_stage.ensure((a1 - a0) * G * cs->cpb * 8);
      HIPCHECK(launch_gather_desc(cs_desc(cs, in).as<uint64_t>(), cs->bmap.as<uint32_t>() + a0 * G, (a1 - a0) * G,
                                  cs->cpb, cs->desc_stage.as<uint64_t>(), stream()));
      rs_encode(cs, out, est, [&] {
        PassDyn dy = cs_dyn(cs, out);
        dy.desc_staged = 1;
        dy.desc_in = cs->desc_stage.as<uint64_t>();
        if (P.passes.size() == 1) {
          dy.bmap = cs->bmap.as<uint32_t>() + a0 * G;
          dy.slot_region = cs->slot_region.as<uint32_t>() + a0;
          HIPCHECK(launch_program_pass(P, 0, IN_DEC, OUT_ENC, dy, a1 - a0, stream()));
          cs_mark(cs, T_GATE);
        } else {
          const uint64_t gpb = std::max<uint64_t>(1, cs->scratch_amps >> L);
          dy.raw = cs->scratch.as<double>();
          for (uint64_t b0 = a0; b0 < a1; b0 += gpb) {
            const uint64_t cnt = std::min(gpb, a1 - b0);
            dy.bmap = cs->bmap.as<uint32_t>() + b0 * G;
            dy.desc_in = cs->desc_stage.as<uint64_t>() + (b0 - a0) * G * cs->cpb;
            dy.slot_region = cs->slot_region.as<uint32_t>() + b0;
            for (size_t p = 0; p < P.passes.size(); ++p) {
              const int im = p == 0 ? IN_DEC : IN_RAW;
              const int om = p + 1 == P.passes.size() ? OUT_ENC : OUT_RAW;
              HIPCHECK(launch_program_pass(P, p, im, om, dy, cnt, stream()));
              cs_mark(cs, pass_timer(im, om));
            }
          }
        }
      }, release_pending, &dem);
      uint64_t dsum = 0;
      for (auto d : dem) dsum += d;
      if (ein > 0) ratio = std::min(64.0, std::max(1.0, (double)dsum / (double)ein)) * 1.1;
    }
    for (uint32_t r = 0; r < last_use.size(); ++r)
      if (last_use[r] == (int64_t)j) pending.push_back(r);
  }
  HIPCHECK(hipEventRecord(cs->ev1, stream()));
  release_pending();
  rs_release_half(cs, in);
  float ms = 0.f;
  HIPCHECK(hipEventSynchronize(cs->ev1));
  HIPCHECK(hipEventElapsedTime(&ms, cs->ev0, cs->ev1));
  cs_collect_marks(cs);
  cs_commit(cs, out);
  const uint64_t bytes_out = rs_current_bytes(cs);
  if (const char* e = std::getenv("QOMPSIM_REGION_LOG"); e && *e == '1') {
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - tw0).count();
    std::fprintf(stderr,
                 "[regions] stage %u: %u batches, %llu launches, %llu retries, map %llu units %.1f ms, unmap %llu "
                 "units %.1f ms, wait %.1f ms, events %.1f ms, wall %.1f ms, mapped %.2f GiB\n",
                 si, nbatch, (unsigned long long)R.n_launch, (unsigned long long)R.n_retry,
                 (unsigned long long)R.n_map, R.t_map * 1e3, (unsigned long long)R.n_unmap, R.t_unmap * 1e3,
                 R.t_wait * 1e3, ms, wall * 1e3, R.mapped_bytes / 1073741824.0);
  }
  cs->cur_bytes = bytes_out;
  cs->bytes_hist.push_back(bytes_out);
  cs->stats.push_back({ms * 1e-3, (double)bytes_in, (double)bytes_out, (double)(nact << L), (double)P.passes.size()});
}

static void cs_register(qsim_cs* cs, std::vector<HGate> gs, const std::vector<std::vector<uint32_t>>& stage_ids,
                        const std::vector<std::vector<uint32_t>>* user_b2g, const std::vector<uint32_t>* user_ng) {
  check_qubits(gs, cs->n);
  cs->gates = std::move(gs);
  std::vector<CStage> st(stage_ids.size());
  for (size_t s = 0; s < stage_ids.size(); ++s) {
    for (auto g : stage_ids[s])
      if (g >= cs->gates.size()) throw std::out_of_range("stage gate id out of range");
    st[s].gate_ids = stage_ids[s];
    stage_groups(cs, st[s]);
    const auto canon = canonical_b2g(cs, st[s]);
    if (user_b2g) {
      // each registered group must be a union of canonical groups
      const auto& ub = (*user_b2g)[s];
      if (ub.size() != cs->nb) throw std::invalid_argument("block_to_group size mismatch");
      std::vector<int64_t> owner(1ull << st[s].O.size(), -1);
      for (uint64_t b = 0; b < cs->nb; ++b) {
        int64_t& o = owner[canon[b]];
        if (o < 0) o = ub[b];
        else if (o != (int64_t)ub[b]) throw std::out_of_range("block group splits a gate pair (block not found in group)");
      }
      st[s].b2g = ub;
      st[s].num_groups = (*user_ng)[s];
    } else {
      st[s].b2g = canon;
      st[s].num_groups = 1u << st[s].O.size();
    }
    build_stage_program(cs, st[s]);
  }
  cs->stages = std::move(st);
}

}  // namespace qsim

// ================================================================= C ABI
#define API_BEGIN try {
#define API_END                                  \
  return 0;                                      \
  }                                              \
  catch (const std::invalid_argument& e) {       \
    g_err = e.what();                            \
    return -2;                                   \
  }                                              \
  catch (const std::out_of_range& e) {           \
    g_err = e.what();                            \
    return -3;                                   \
  }                                              \
  catch (const std::domain_error& e) {           \
    g_err = e.what();                            \
    return -2;                                   \
  }                                              \
  catch (const std::exception& e) {              \
    g_err = e.what();                            \
    return -1;                                   \
  }                                              \
  catch (...) {                                  \
    g_err = "unknown error";                     \
    return -1;                                   \
  }

static void require(bool c, const char* msg) {
  if (!c) throw std::invalid_argument(msg);
}

extern "C" {

const char* qsim_last_error(void) { return g_err.c_str(); }
int qsim_version(void) { return 1; }

int qsim_device_count(int* count) {
  API_BEGIN
  HIPCHECK(hipGetDeviceCount(count));
  API_END
}

int qsim_synchronize(void) {
  API_BEGIN
  HIPCHECK(hipStreamSynchronize(stream()));
  API_END
}

uint64_t qsim_optimize_pool_bound(uint64_t n_gates, uint64_t pool_len, uint32_t max_block_size) {
  const uint64_t d = 1ull << std::min<uint32_t>(max_block_size, 12);
  return pool_len + n_gates * d * d * 2 + 16;
}

static void write_gates(const std::vector<HGate>& gs, qsim_gate* out, double* pool, uint64_t cap) {
  uint64_t off = 0;
  for (size_t i = 0; i < gs.size(); ++i) {
    const HGate& g = gs[i];
    qsim_gate q{};
    q.n_targets = (uint32_t)g.t.size();
    q.n_controls = (uint32_t)g.c.size();
    for (size_t j = 0; j < g.t.size(); ++j) {
      q.targets[j] = g.t[j];
      q.target_commute[j] = g.tprop[j];
      q.pauli_ids[j] = j < g.pauli.size() ? g.pauli[j] : 0;
    }
    for (size_t j = 0; j < g.c.size(); ++j) {
      q.controls[j] = g.c[j];
      q.control_values[j] = g.cval[j];
    }
    q.flags = g.flags;
    q.param_index = g.param;
    q.matrix_offset = off;
    if (2 * (off + g.m.size()) > cap) throw std::length_error("optimizer output pool too small");
    for (size_t k = 0; k < g.m.size(); ++k) {
      pool[2 * (off + k)] = g.m[k].real();
      pool[2 * (off + k) + 1] = g.m[k].imag();
    }
    off += g.m.size();
    out[i] = q;
  }
}

int qsim_optimize(const qsim_gate* gates, uint64_t n_gates, const double* pool, uint64_t pool_len,
                  uint32_t max_block_size, qsim_gate* out, uint64_t* n_out, double* pool_out, uint64_t* pool_out_len,
                  int64_t* origin) {
  API_BEGIN
  auto gs = parse_gates(gates, n_gates, pool, pool_len);
  auto opt = optimize_circuit(std::move(gs), max_block_size);
  uint64_t need = 0;
  for (auto& g : opt) need += 2 * g.m.size();
  require(need <= *pool_out_len, "qsim_optimize: pool_out too small");
  write_gates(opt, out, pool_out, *pool_out_len);
  if (origin)
    for (size_t i = 0; i < opt.size(); ++i) origin[i] = opt[i].origin;
  *n_out = opt.size();
  *pool_out_len = need;
  API_END
}

// ---------------------------------------------------------------- SV API
int qsim_sv_create(uint32_t n_qubits, qsim_sv** out) {
  API_BEGIN
  require(n_qubits >= 1 && n_qubits <= 40, "qubit count out of range");
  auto sv = std::make_unique<qsim_sv>();
  sv->n = n_qubits;
  sv->dim = 1ull << n_qubits;
  sv->d.alloc(sv->dim * 16);
  HIPCHECK(launch_fill_basis(sv->ptr(), sv->dim, 0, stream()));
  HIPCHECK(hipStreamSynchronize(stream()));
  *out = sv.release();
  API_END
}

int qsim_sv_destroy(qsim_sv* sv) {
  API_BEGIN
  delete sv;
  API_END
}

int qsim_sv_set_basis(qsim_sv* sv, uint64_t index) {
  API_BEGIN
  require(index < sv->dim, "basis index out of range");
  HIPCHECK(launch_fill_basis(sv->ptr(), sv->dim, index, stream()));
  HIPCHECK(hipStreamSynchronize(stream()));
  API_END
}

int qsim_sv_load(qsim_sv* sv, const double* host) {
  API_BEGIN
  HIPCHECK(hipMemcpy(sv->d.p, host, sv->dim * 16, hipMemcpyHostToDevice));
  API_END
}

int qsim_sv_copy(qsim_sv* dst, const qsim_sv* src) {
  API_BEGIN
  require(dst->n == src->n, "qubit count mismatch");
  HIPCHECK(hipMemcpyAsync(dst->d.p, src->d.p, src->dim * 16, hipMemcpyDeviceToDevice, stream()));
  HIPCHECK(hipStreamSynchronize(stream()));
  API_END
}

int qsim_sv_read(const qsim_sv* sv, double* host) {
  API_BEGIN
  HIPCHECK(hipStreamSynchronize(stream()));
  HIPCHECK(hipMemcpy(host, sv->d.p, sv->dim * 16, hipMemcpyDeviceToHost));
  API_END
}

int qsim_sv_apply(qsim_sv* sv, const qsim_gate* gates, uint64_t n_gates, const double* pool, uint64_t pool_len) {
  API_BEGIN
  auto gs = parse_gates(gates, n_gates, pool, pool_len);
  check_qubits(gs, sv->n);
  sv_apply_gates(sv, gs);
  API_END
}

int qsim_sv_norm_squared(const qsim_sv* sv, double* out) {
  API_BEGIN
  const unsigned grid = reduce_grid(sv->dim);
  sv->part.ensure(1, grid);
  HIPCHECK(launch_sv_norm(sv->ptr(), sv->dim, sv->part.d.as<double>(), grid, stream()));
  sv->part.fetch(1, grid);
  DDAcc acc;
  sv->part.accumulate(1, grid, &acc);
  *out = acc.value();
  API_END
}

int qsim_sv_normalize(qsim_sv* sv, double squared_norm) {
  API_BEGIN
  require(squared_norm > 0.0, "normalize: non-positive norm");
  HIPCHECK(launch_scale(sv->ptr(), sv->dim, std::sqrt(1.0 / squared_norm), stream()));
  HIPCHECK(hipStreamSynchronize(stream()));
  API_END
}

int qsim_sv_scale(qsim_sv* sv, double factor) {
  API_BEGIN
  HIPCHECK(launch_scale(sv->ptr(), sv->dim, factor, stream()));
  HIPCHECK(hipStreamSynchronize(stream()));
  API_END
}

int qsim_sv_inner_product(const qsim_sv* bra, const qsim_sv* ket, double* re, double* im) {
  API_BEGIN
  require(bra->n == ket->n, "qubit count mismatch");
  const unsigned grid = reduce_grid(bra->dim);
  bra->part.ensure(2, grid);
  HIPCHECK(launch_sv_inner(bra->ptr(), ket->ptr(), bra->dim, bra->part.d.as<double>(), grid, stream()));
  bra->part.fetch(2, grid);
  DDAcc acc[2];
  bra->part.accumulate(2, grid, acc);
  *re = acc[0].value();
  *im = acc[1].value();
  API_END
}

static void ipow(uint32_t ny, double& r, double& i) {
  switch (ny & 3u) {
    case 0: r = 1; i = 0; break;
    case 1: r = 0; i = 1; break;
    case 2: r = -1; i = 0; break;
    default: r = 0; i = -1; break;
  }
}


namespace qsim {
// Pauli-sum tables: terms grouped by flip mask (X/Y qubits); the coefficient
// of a term is scale * c * i^ny.
static uint32_t build_pauli_groups(uint64_t n_terms, const uint64_t* xm, const uint64_t* zm, const uint32_t* ny,
                                   const double* cre, const double* cim, double scale, std::vector<PauliGroup>& G,
                                   std::vector<PauliTermZ>& T) {
  std::map<uint64_t, std::vector<uint64_t>> by_flip;
  for (uint64_t t = 0; t < n_terms; ++t) by_flip[xm[t]].push_back(t);
  G.clear();
  T.clear();
  for (auto& kv : by_flip) {
    PauliGroup g{};
    g.flip = kv.first;
    g.t0 = (uint32_t)T.size();
    for (auto t : kv.second) {
      double pr, pi;
      ipow(ny[t], pr, pi);
      PauliTermZ z{};
      z.z = zm[t];
      z.cr = scale * (cre[t] * pr - cim[t] * pi);
      z.ci = scale * (cre[t] * pi + cim[t] * pr);
      T.push_back(z);
    }
    g.nt = (uint32_t)(T.size() - g.t0);
    G.push_back(g);
  }
  return (uint32_t)G.size();
}

}  // namespace qsim

int qsim_sv_expectation_pauli(const qsim_sv* sv, uint64_t n_terms, const uint64_t* xm, const uint64_t* zm,
                              const uint32_t* ny, const double* cre, const double* cim, double* ore, double* oim) {
  API_BEGIN
  require(sv != nullptr && ore != nullptr && oim != nullptr, "null argument");
  if (n_terms == 0) {
    *ore = *oim = 0.0;
    return 0;
  }
  // every term in one sweep: terms sharing an X/Y flip mask share the load of
  // the flipped amplitude
  std::vector<PauliGroup> G;
  std::vector<PauliTermZ> T;
  const uint32_t ng = build_pauli_groups(n_terms, xm, zm, ny, cre, cim, 1.0, G, T);
  qsim_sv* s = const_cast<qsim_sv*>(sv);
  upload(s->adj_groups, G.data(), G.size() * sizeof(PauliGroup));
  upload(s->adj_terms, T.data(), T.size() * sizeof(PauliTermZ));
  const unsigned grid = reduce_grid(sv->dim);
  s->part.ensure(2, grid);  // (re, im) compensated pairs per block
  HIPCHECK(launch_pauli_expect(sv->ptr(), sv->dim, s->adj_groups.as<PauliGroup>(), ng, s->adj_terms.as<PauliTermZ>(),
                               s->part.d.as<double>(), grid, stream()));
  s->part.fetch(2, grid);
  DDAcc acc[2];
  s->part.accumulate(2, grid, acc);
  *ore = acc[0].value();
  *oim = acc[1].value();
  API_END
}

// The same expectation with every device buffer in the library's static
// scratch array: term tables, block partials and final sums.  The grouping
// (flip masks ascending, terms in input order within a group) and the
// reduction order are those of qsim_sv_expectation_pauli, so the results are
// equal bit for bit.  No heap or device allocation: the tables are built in
// fixed-size arrays on the stack.
int qsim_sv_expectation_pauli_noalloc(const qsim_sv* sv, uint64_t n_terms, const uint64_t* xm, const uint64_t* zm,
                                      const uint32_t* ny, const double* cre, const double* cim, double* ore,
                                      double* oim) {
  API_BEGIN
  require(sv != nullptr && ore != nullptr && oim != nullptr, "null argument");
  if (n_terms > kNoAllocMaxTerms)
    throw std::invalid_argument("qsim_sv_expectation_pauli_noalloc: at most " + std::to_string(kNoAllocMaxTerms) +
                                " terms (got " + std::to_string(n_terms) + ")");
  if (n_terms == 0) {
    *ore = *oim = 0.0;
    return 0;
  }
  require(xm != nullptr && zm != nullptr && ny != nullptr && cre != nullptr && cim != nullptr, "null argument");
  uint32_t order[kNoAllocMaxTerms];
  for (uint32_t t = 0; t < (uint32_t)n_terms; ++t) order[t] = t;
  std::sort(order, order + n_terms, [&](uint32_t a, uint32_t b) { return xm[a] != xm[b] ? xm[a] < xm[b] : a < b; });
  PauliGroup G[kNoAllocMaxTerms];
  PauliTermZ T[kNoAllocMaxTerms];
  uint32_t ng = 0;
  for (uint32_t i = 0; i < (uint32_t)n_terms; ++i) {
    const uint32_t t = order[i];
    if (i == 0 || xm[t] != G[ng - 1].flip) {
      G[ng] = PauliGroup{};
      G[ng].flip = xm[t];
      G[ng].t0 = i;
      ++ng;
    }
    ++G[ng - 1].nt;
    double pr, pi;
    ipow(ny[t], pr, pi);
    T[i].z = zm[t];
    T[i].cr = 1.0 * (cre[t] * pr - cim[t] * pi);
    T[i].ci = 1.0 * (cre[t] * pi + cim[t] * pr);
  }
  void* scr = nullptr;
  HIPCHECK(noalloc_scratch(&scr));
  unsigned char* sb = static_cast<unsigned char*>(scr);
  const unsigned grid = reduce_grid(sv->dim);
  if (grid > kNoAllocMaxGrid) throw std::logic_error("reduction grid exceeds the scratch partials");
  HIPCHECK(hipMemcpyAsync(sb, G, (size_t)ng * sizeof(PauliGroup), hipMemcpyHostToDevice, stream()));
  HIPCHECK(hipMemcpyAsync(sb + kNoAllocTermsOff, T, (size_t)n_terms * sizeof(PauliTermZ), hipMemcpyHostToDevice,
                          stream()));
  HIPCHECK(launch_pauli_expect_noalloc(sv->ptr(), sv->dim, ng, sb, grid, stream()));
  double fin[4];
  HIPCHECK(hipMemcpyAsync(fin, sb + kNoAllocFinalOff, sizeof(fin), hipMemcpyDeviceToHost, stream()));
  HIPCHECK(hipStreamSynchronize(stream()));
  *ore = fin[0] + fin[1];
  *oim = fin[2] + fin[3];
  API_END
}

namespace qsim {

static HGate inverse_gate(const HGate& g) {
  HGate h = g;
  const size_t D = 1ull << g.t.size();
  for (size_t r = 0; r < D; ++r)
    for (size_t c = 0; c < D; ++c) h.m[c * D + r] = std::conj(g.m[r * D + c]);
  h.flags &= ~QSIM_GATE_PARAMETRIC;
  return h;
}

// Gradient point of a parametric Pauli rotation R_P(theta) = exp(+i theta P/2).
static LGate grad_point(const HGate& g) {
  LGate L;
  L.grad = g.param;
  for (size_t i = 0; i < g.t.size(); ++i) {
    const uint64_t bit = 1ull << g.t[i];
    const uint32_t pid = g.pauli[i];
    if (pid == 1 || pid == 2) L.gx |= bit;
    if (pid == 2 || pid == 3) L.gz |= bit;
    if (pid == 2) L.nY++;
  }
  for (uint32_t q = 0; q < 64; ++q) {
    const uint64_t bit = 1ull << q;
    if (L.gx & bit) L.N.push_back(q);
    else if (L.gz & bit) (L.gx ? L.S : L.Q).push_back(q);
  }
  L.diag = L.gx == 0;
  return L;
}

// Reverse programs run on the wave-tile engine when its shape fits (at least
// 10 qubits, gates and gradient points on at most two X/Y qubits;
// QOMPSIM_ADJ_ENGINE=lds forces the LDS-tile engine), else on LDS tiles of
// 2^tb amplitudes.
static bool adjoint_wave_ok(const std::vector<LGate>& lg, uint32_t n) {
  const char* e = std::getenv("QOMPSIM_ADJ_ENGINE");
  if (e && std::string(e) == "lds") return false;
  return wave_engine_ok(lg, n);
}

// Wave forms: by default a pair of waves per 1024-amplitude tile (psi and
// lambda in one wave each, the forward engine's tile).  QOMPSIM_ADJ_ENGINE=pair9
// pairs waves of 8 rows on 512-amplitude tiles (three waves per SIMD);
// QOMPSIM_ADJ_ENGINE=wave9 selects one wave holding 8 rows of each state.
static uint32_t adjoint_wave_tile() {
  const char* e = std::getenv("QOMPSIM_ADJ_ENGINE");
  return (e && (std::string(e) == "wave9" || std::string(e) == "pair9")) ? 9u : kWaveTile;
}
static bool adjoint_pair9() {
  const char* e = std::getenv("QOMPSIM_ADJ_ENGINE");
  return e && std::string(e) == "pair9";
}
constexpr uint32_t kAdjWaveTile = kWaveTile;

// A gradient point reads psi and lambda without changing them, and a gate
// that commutes with the point's Pauli string P leaves <lambda|P|psi> as it is
// when applied to both states, so a point may move back over such gates and
// over other points.  Each point can sit anywhere from the first earlier gate
// it does not commute with up to its own place; the points are gathered at as
// few places as possible (greedy stabbing of those intervals, taken by their
// right ends), so that runs of adjacent points share one exchange of the psi
// tile in the paired kernel.  Commutation uses the scheduler's mask test.
// QOMPSIM_ADJ_CLUSTER=0 keeps the order [GRAD(i), inverse(g_i)].
static std::vector<LGate> cluster_gradient_points(const std::vector<LGate>& lg) {
  const char* e = std::getenv("QOMPSIM_ADJ_CLUSTER");
  if (e && *e == '0') return lg;
  const size_t n = lg.size();
  auto qmask = [](const std::vector<uint32_t>& v) {
    uint64_t m = 0;
    for (auto q : v) m |= 1ull << q;
    return m;
  };
  std::vector<uint64_t> ndm(n), allm(n);
  for (size_t i = 0; i < n; ++i) {
    ndm[i] = lg[i].diag ? 0ull : qmask(lg[i].N);
    allm[i] = lg[i].diag ? qmask(lg[i].Q) : (qmask(lg[i].N) | qmask(lg[i].S));
  }
  std::vector<size_t> left(n, 0), points;
  for (size_t i = 0; i < n; ++i) {
    if (lg[i].grad < 0) continue;
    points.push_back(i);
    size_t j = i;
    while (j > 0) {
      const size_t k = j - 1;
      if (lg[k].grad < 0 && ((ndm[k] & allm[i]) || (allm[k] & ndm[i]))) break;
      j = k;
    }
    left[i] = j;
  }
  // anchor[i]: the place point i moves to (the place of its cluster's first point)
  std::vector<size_t> anchor(n, SIZE_MAX);
  for (size_t a = 0; a < points.size(); ++a) {
    const size_t p = points[a];
    if (anchor[p] != SIZE_MAX) continue;
    for (size_t b = a; b < points.size(); ++b)
      if (anchor[points[b]] == SIZE_MAX && left[points[b]] <= p) anchor[points[b]] = p;
  }
  std::vector<LGate> out;
  out.reserve(n);
  for (size_t i = 0; i < n; ++i) {
    if (lg[i].grad < 0) {
      out.push_back(lg[i]);
    } else if (anchor[i] == i) {
      for (size_t b : points)
        if (anchor[b] == i) out.push_back(lg[b]);
    }
  }
  return out;
}

// Moves each diagonal op of a reverse program back over earlier ops it
// commutes with (other diagonal ops; dense ops whose non-diagonal targets
// avoid its qubits; gradient points whose X/Y qubits avoid them), so that
// the diagonal gates of a layer (HEA's RZ^-1 and CZ^-1) gather in one place
// instead of one per qubit between the points.
static std::vector<LGate> gather_reverse_diagonals(const std::vector<LGate>& lg) {
  auto qmask = [](const std::vector<uint32_t>& v) {
    uint64_t m = 0;
    for (auto q : v) m |= 1ull << q;
    return m;
  };
  std::vector<LGate> out;
  std::vector<uint64_t> blk;  // per output op: the qubits a diagonal op may not cross
  out.reserve(lg.size());
  blk.reserve(lg.size());
  for (const LGate& g : lg) {
    const bool dg = g.grad < 0 && g.diag;
    uint64_t b = 0;
    if (g.grad >= 0) b = g.gx;
    else if (!g.diag) b = qmask(g.N);
    size_t at = out.size();
    if (dg) {
      const uint64_t q = qmask(g.Q);
      while (at > 0 && !(blk[at - 1] & q)) --at;
    }
    out.insert(out.begin() + (ptrdiff_t)at, g);
    blk.insert(blk.begin() + (ptrdiff_t)at, b);
  }
  return out;
}

// wave_tile: 0 for the default wave form (adjoint_wave_tile), else the tile
// bits of the wave passes
static void plan_reverse(Program& P, const std::vector<LGate>& lg0, uint32_t n, uint32_t tb, bool wave,
                         uint32_t wave_tile = 0) {
  const std::vector<LGate> lg = cluster_gradient_points(gather_reverse_diagonals(lg0));
  const uint32_t wt = wave_tile ? wave_tile : adjoint_wave_tile();
  if (wave && n >= wt) plan_program_engine(P, lg, n, wt, 6, false, false, true, wt);
  else plan_program_engine(P, lg, n, tb, std::min<uint32_t>(5, tb), false, false, false);
}

// tiles of pass p of a reverse program over 2^n amplitudes
static uint64_t reverse_tiles(const Program& P, size_t p, uint32_t n) { return (1ull << n) >> P.passes[p].t; }

static hipError_t launch_reverse_pass(const Program& P, size_t p, uint32_t n, double2* psi, double2* lam,
                                      double* partials, hipStream_t st) {
  const uint64_t ntiles = reverse_tiles(P, p, n);
  if (P.passes[p].engine == ENGINE_WAVE) {
    if (P.passes[p].t == kWaveTile || adjoint_pair9())
      return launch_wave_pair_adjoint((int)P.passes[p].t, P.dargs(p), psi, lam, partials, ntiles, st);
    return launch_wave_adjoint((int)P.passes[p].t, P.dargs(p), psi, lam, partials, ntiles, st);
  }
  return launch_adjoint_pass((int)P.passes[p].t, P.dargs(p), psi, lam, partials, ntiles, st);
}

static uint32_t adjoint_tile_bits(uint32_t n) {
  uint32_t tb = 10;
  if (const char* e = std::getenv("QOMPSIM_ADJ_TILE_BITS")) tb = (uint32_t)std::atoi(e);
  tb = std::min<uint32_t>(std::max<uint32_t>(tb, 4), 12);
  return std::min(tb, n);
}
}  // namespace qsim

namespace qsim {
// Gates the adjoint accepts (ParametricQuantumCircuit::backprop rejects other
// parametric gates): parametric gates are Pauli rotations with a valid
// parameter index; every gate has at most 3 non-diagonal targets.
static void validate_backprop_gates(const std::vector<HGate>& gs, uint64_t n_params) {
  for (size_t i = 0; i < gs.size(); ++i) {
    const HGate& g = gs[i];
    if (!g.parametric()) continue;
    if (g.param < 0 || (uint64_t)g.param >= n_params)
      throw std::out_of_range("gate " + std::to_string(i) + ": parameter index out of range");
    bool pauli_ok = !g.c.empty() ? false : !g.pauli.empty();
    for (auto pid : g.pauli) pauli_ok = pauli_ok && pid >= 1 && pid <= 3;
    if (!pauli_ok)
      throw std::invalid_argument("gate " + std::to_string(i) + " does not support backprop (not a Pauli rotation)");
    if (g.t.size() > 3) {
      uint32_t nxy = 0;
      for (auto pid : g.pauli) nxy += pid != 3;
      if (nxy > 3) throw std::invalid_argument("backprop: Pauli rotation with more than 3 X/Y factors");
    }
  }
  for (auto& g : gs) {
    uint32_t nd = 0;
    const size_t K = g.t.size(), D = 1ull << K;
    for (size_t i = 0; i < K; ++i) {
      bool dq = true;
      for (size_t r = 0; r < D && dq; ++r)
        for (size_t c = 0; c < D; ++c)
          if ((((r ^ c) >> i) & 1u) && g.m[r * D + c] != cplx(0, 0)) { dq = false; break; }
      nd += !dq;
    }
    if (nd > 3) throw std::invalid_argument("backprop: gates with more than 3 non-diagonal targets");
  }
}

// lambda = 2 H psi (one sweep over Pauli groups); returns <H> = Re<psi|lambda>/2
static double lambda_and_energy(qsim_sv* sv, const double2* psi, double2* lam, uint64_t n_terms, const uint64_t* xm,
                                const uint64_t* zm, const uint32_t* ny, const double* cre, const double* cim) {
  std::vector<PauliGroup> G;
  std::vector<PauliTermZ> T;
  const uint32_t ng = build_pauli_groups(n_terms, xm, zm, ny, cre, cim, 2.0, G, T);
  upload(sv->adj_groups, G.data(), G.size() * sizeof(PauliGroup));
  upload(sv->adj_terms, T.data(), T.size() * sizeof(PauliTermZ));
  const unsigned grid = reduce_grid(sv->dim);
  sv->part.ensure(1, grid);
  HIPCHECK(launch_pauli_sum(psi, lam, sv->dim, sv->adj_groups.as<PauliGroup>(), ng, sv->adj_terms.as<PauliTermZ>(),
                            sv->part.d.as<double>(), grid, stream()));
  sv->part.fetch(1, grid);
  DDAcc eacc;
  sv->part.accumulate(1, grid, &eacc);
  return eacc.value() / 2.0;
}

// Reverse sweep over gates gs[b, e) of one state pair: psi_w holds the state
// after gate e-1 and lambda the costate there; on return both are pulled back
// to before gate b and the gradient points of the range are added to gacc.
static void reverse_sweep(qsim_sv* sv, double2* psi_w, double2* lam, const std::vector<HGate>& gs, size_t b, size_t e,
                          std::vector<DDAcc>& gacc) {
  std::vector<int> qmap(sv->n);
  std::iota(qmap.begin(), qmap.end(), 0);
  std::vector<LGate> lg;
  for (size_t i = e; i-- > b;) {
    if (gs[i].parametric()) lg.push_back(grad_point(gs[i]));
    lg.push_back(lower_gate(inverse_gate(gs[i]), qmap));
  }
  if (lg.empty()) return;
  const uint32_t tb = adjoint_tile_bits(sv->n);
  Program P;
  plan_reverse(P, lg, sv->n, tb, adjoint_wave_ok(lg, sv->n));
  // renumber gradient points per pass
  std::vector<std::vector<int32_t>> params(P.passes.size());
  uint32_t maxpts = 0;
  for (size_t pi = 0; pi < P.passes.size(); ++pi) {
    for (uint32_t o = P.passes[pi].op_begin; o < P.passes[pi].op_end; ++o)
      if (P.ops[o].kind == OP_GRAD) {
        params[pi].push_back((int32_t)P.ops[o].gidx);
        P.ops[o].gidx = (uint32_t)(params[pi].size() - 1);
      }
    maxpts = std::max<uint32_t>(maxpts, (uint32_t)params[pi].size());
  }
  P.upload_all();
  uint64_t maxtiles = 1;
  for (size_t pi = 0; pi < P.passes.size(); ++pi) maxtiles = std::max(maxtiles, reverse_tiles(P, pi, sv->n));
  // each pass's reduced sums go to their own slots of adj_out; one copy and
  // one wait after the last pass, then the merges in pass order
  std::vector<uint64_t> out_off(P.passes.size() + 1, 0);
  for (size_t pi = 0; pi < P.passes.size(); ++pi) out_off[pi + 1] = out_off[pi] + params[pi].size();
  const uint64_t nout = out_off.back();
  sv->adj_part.ensure(std::max<uint64_t>(1, (uint64_t)maxpts * maxtiles) * 2 * sizeof(double));
  sv->adj_out.ensure(std::max<uint64_t>(1, nout) * 2 * sizeof(double));
  sv->adj_stage.ensure((uint64_t)std::max<uint32_t>(1, maxpts) * kGradReduceChunks * 2 * sizeof(double));
  for (size_t pi = 0; pi < P.passes.size(); ++pi) {
    HIPCHECK(launch_reverse_pass(P, pi, sv->n, psi_w, lam, sv->adj_part.as<double>(), stream()));
    const uint32_t np = (uint32_t)params[pi].size();
    if (np == 0) continue;
    HIPCHECK(launch_grad_reduce(sv->adj_part.as<double>(), reverse_tiles(P, pi, sv->n), np,
                                sv->adj_out.as<double>() + 2 * out_off[pi], sv->adj_stage.as<double>(), stream()));
  }
  if (nout == 0) return;
  std::vector<double> hout(2 * nout);
  HIPCHECK(hipMemcpyAsync(hout.data(), sv->adj_out.p, nout * 2 * sizeof(double), hipMemcpyDeviceToHost, stream()));
  HIPCHECK(hipStreamSynchronize(stream()));
  for (size_t pi = 0; pi < P.passes.size(); ++pi)
    for (uint32_t k = 0; k < (uint32_t)params[pi].size(); ++k)
      gacc[params[pi][k]].merge(hout[2 * (out_off[pi] + k)], hout[2 * (out_off[pi] + k) + 1]);
}

// ------------------------------------------------ lambda-only reverse sweep
// QOMPSIM_ADJ_ENGINE=lam, when device memory allows.  The reverse program is
// planned so that every gradient point sits at the start of a pass; the
// forward pass runs the same passes (gates un-inverted, in reverse order) and
// keeps psi at the start of every reverse pass that has points; the reverse
// sweep then carries lambda alone and reads psi from those copies.  Points at
// pass starts cut the passes at every run of points (HEA-28: 49 passes
// instead of 23, and 48 copies of psi), so the paired sweep stays the default.
static bool adjoint_lam_engine() {
  const char* e = std::getenv("QOMPSIM_ADJ_ENGINE");
  return e && std::string(e) == "lam";
}

static bool backprop_lam(qsim_sv* sv, const std::vector<HGate>& gs, uint64_t n_terms, const uint64_t* xm,
                         const uint64_t* zm, const uint32_t* ny, const double* cre, const double* cim,
                         std::vector<DDAcc>& gacc, double* energy) {
  const uint32_t n = sv->n;
  if (!adjoint_lam_engine() || n < kWaveTile || gs.empty()) return false;
  std::vector<int> qmap(n);
  std::iota(qmap.begin(), qmap.end(), 0);
  std::vector<LGate> lg;
  for (size_t i = gs.size(); i-- > 0;) {
    if (gs[i].parametric()) lg.push_back(grad_point(gs[i]));
    lg.push_back(lower_gate(inverse_gate(gs[i]), qmap));
    lg.back().orig = (int64_t)i;
  }
  if (!wave_engine_ok(lg, n)) return false;
  lg = cluster_gradient_points(gather_reverse_diagonals(lg));
  // reverse program: one planning call per segment starting at a run of
  // points, so that points land at pass starts (a point the planner defers
  // is taken first by a later pass of its segment)
  Program R;
  std::vector<std::vector<size_t>> rgates;  // per reverse pass: indices into lg
  {
    size_t b = 0;
    while (b < lg.size()) {
      size_t e = b + 1;
      while (e < lg.size() && !(lg[e].grad >= 0 && lg[e - 1].grad < 0)) ++e;
      std::vector<LGate> seg(lg.begin() + (ptrdiff_t)b, lg.begin() + (ptrdiff_t)e);
      std::vector<std::vector<size_t>> pg;
      plan_program_engine(R, seg, n, kWaveTile, 6, false, false, true, kWaveTile, &pg);
      for (auto& v : pg) {
        for (auto& i : v) i += b;
        rgates.push_back(v);
      }
      b = e;
    }
  }
  const size_t m = R.passes.size();
  if (m == 0) return false;
  std::vector<char> has_pts(m, 0);
  for (size_t r = 0; r < m; ++r) {
    const PassArgs& a = R.passes[r];
    if (a.engine != ENGINE_WAVE || a.t != kWaveTile) return false;
    bool lead = true;
    for (uint32_t o = a.op_begin; o < a.op_end; ++o) {
      if (R.ops[o].kind == OP_GRAD) {
        if (!lead) return false;  // a point inside a pass: the paired engine takes the circuit
        has_pts[r] = 1;
      } else {
        lead = false;
      }
    }
  }
  // device memory: a copy of psi per reverse pass with points (pass 0 reads
  // the state itself), a work copy, lambda
  const uint64_t bytes = sv->dim * sizeof(double2);
  uint64_t nck = 0;
  for (size_t r = 1; r < m; ++r) nck += has_pts[r];
  {
    size_t free_b = 0, total_b = 0;
    HIPCHECK(hipMemGetInfo(&free_b, &total_b));
    const uint64_t held = sv->adj_ck_bytes() + sv->adj_lam.bytes + sv->adj_psi.bytes;
    const uint64_t want = (nck + 2) * bytes;
    if (want > held + free_b - std::min<uint64_t>(free_b, 1ull << 30)) return false;
  }
  // forward: reverse pass r's gates un-inverted, in reverse order, for r =
  // m-1 .. 0; the output of segment r is psi at the start of reverse pass r
  Program F;
  std::vector<size_t> fbeg(m, 0), fend(m, 0);  // segment r: F passes [fbeg[r], fend[r])
  for (size_t r = m; r-- > 0;) {
    std::vector<LGate> fg;
    for (size_t k = rgates[r].size(); k-- > 0;) {
      const LGate& g = lg[rgates[r][k]];
      if (g.grad >= 0) continue;
      fg.push_back(lower_gate(gs[(size_t)g.orig], qmap));
    }
    fbeg[r] = F.passes.size();
    if (!fg.empty()) plan_program_engine(F, fg, n, default_tile_bits(), 5, false, false, wave_engine_ok(fg, n));
    fend[r] = F.passes.size();
  }
  if (std::getenv("QOMPSIM_PLAN_LOG"))
    fprintf(stderr, "[adjoint] lambda-only sweep: %zu reverse passes, %zu forward passes, %llu copies of psi\n", m,
            F.passes.size(), (unsigned long long)nck);
  while (sv->adj_ck.size() < nck) sv->adj_ck.push_back(std::make_unique<DBuf>());
  for (uint64_t k = 0; k < nck; ++k) sv->adj_ck[k]->ensure(bytes);
  sv->adj_psi.ensure(bytes);  // the work copy
  sv->adj_lam.ensure(bytes);
  std::vector<double2*> ck(m, nullptr);
  {
    size_t k = 0;
    ck[0] = sv->ptr();
    for (size_t r = 1; r < m; ++r)
      if (has_pts[r]) ck[r] = sv->adj_ck[k++]->as<double2>();
  }
  F.upload_all();
  double2* cur = sv->ptr();
  double2* work = sv->adj_psi.as<double2>();
  for (size_t r = m; r-- > 0;) {
    double2* dst = ck[r] ? ck[r] : work;
    if (fbeg[r] == fend[r]) {
      if (dst != cur) HIPCHECK(hipMemcpyAsync(dst, cur, bytes, hipMemcpyDeviceToDevice, stream()));
    } else {
      for (size_t i = fbeg[r]; i < fend[r]; ++i) {
        PassDyn dy{};
        const bool first = i == fbeg[r];
        if (first && dst != cur && F.passes[i].engine != ENGINE_WAVE) {
          HIPCHECK(hipMemcpyAsync(dst, cur, bytes, hipMemcpyDeviceToDevice, stream()));
        }
        const bool oop = first && dst != cur && F.passes[i].engine == ENGINE_WAVE;
        dy.raw = reinterpret_cast<double*>(oop ? cur : dst);
        dy.raw_out = oop ? reinterpret_cast<double*>(dst) : nullptr;
        HIPCHECK(launch_program_pass(F, i, IN_RAW, OUT_RAW, dy, 1, stream()));
      }
    }
    cur = dst;
  }
  // lambda = 2 H psi at the end of the circuit
  *energy = lambda_and_energy(sv, sv->ptr(), sv->adj_lam.as<double2>(), n_terms, xm, zm, ny, cre, cim);
  // reverse: points of pass r against psi at its start, then lambda through
  // the pass's inverted gates
  std::vector<std::vector<int32_t>> params(m);
  uint32_t maxpts = 0;
  for (size_t r = 0; r < m; ++r) {
    for (uint32_t o = R.passes[r].op_begin; o < R.passes[r].op_end; ++o)
      if (R.ops[o].kind == OP_GRAD) {
        params[r].push_back((int32_t)R.ops[o].gidx);
        R.ops[o].gidx = (uint32_t)(params[r].size() - 1);
      }
    maxpts = std::max<uint32_t>(maxpts, (uint32_t)params[r].size());
  }
  R.upload_all();
  const uint64_t ntiles = sv->dim >> kWaveTile;
  sv->adj_part.ensure(std::max<uint64_t>(1, (uint64_t)maxpts * ntiles) * 2 * sizeof(double));
  sv->adj_out.ensure(std::max<uint32_t>(1, maxpts) * 2 * sizeof(double));
  sv->adj_stage.ensure((uint64_t)std::max<uint32_t>(1, maxpts) * kGradReduceChunks * 2 * sizeof(double));
  std::vector<double> hout(2 * std::max<uint32_t>(1, maxpts));
  for (size_t r = 0; r < m; ++r) {
    PassDyn dy{};
    dy.raw = sv->adj_lam.as<double>();
    dy.psi = has_pts[r] ? reinterpret_cast<const double*>(ck[r]) : nullptr;
    dy.partials = sv->adj_part.as<double>();
    dy.part_tiles = ntiles;
    HIPCHECK(launch_wave_lam(R.dargs(r), dy, ntiles, stream()));
    const uint32_t np = (uint32_t)params[r].size();
    if (np == 0) continue;
    HIPCHECK(launch_grad_reduce(sv->adj_part.as<double>(), ntiles, np, sv->adj_out.as<double>(),
                                sv->adj_stage.as<double>(), stream()));
    HIPCHECK(hipMemcpyAsync(hout.data(), sv->adj_out.p, np * 2 * sizeof(double), hipMemcpyDeviceToHost, stream()));
    HIPCHECK(hipStreamSynchronize(stream()));
    for (uint32_t k = 0; k < np; ++k) gacc[params[r][k]].merge(hout[2 * k], hout[2 * k + 1]);
  }
  return true;
}
}  // namespace qsim

int qsim_sv_backprop(qsim_sv* sv, const qsim_gate* gates, uint64_t n_gates, const double* pool, uint64_t pool_len,
                     uint64_t n_terms, const uint64_t* xm, const uint64_t* zm, const uint32_t* ny, const double* cre,
                     const double* cim, uint64_t n_params, double* energy, double* grads) {
  API_BEGIN
  require(sv != nullptr && energy != nullptr && (grads != nullptr || n_params == 0), "null argument");
  require(sv->n >= 4, "backprop needs at least 4 qubits");
  auto gs = parse_gates(gates, n_gates, pool, pool_len);
  check_qubits(gs, sv->n);
  validate_backprop_gates(gs, n_params);
  const uint64_t bytes = sv->dim * sizeof(double2);
  std::vector<DDAcc> gacc(n_params);
  if (!backprop_lam(sv, gs, n_terms, xm, zm, ny, cre, cim, gacc, energy)) {
    // forward: psi = U psi_in
    sv_apply_gates(sv, gs);
    // lambda = 2 H psi; <H> = Re<psi|lambda> / 2
    sv->adj_lam.ensure(bytes);
    sv->adj_psi.ensure(bytes);
    *energy = lambda_and_energy(sv, sv->ptr(), sv->adj_lam.as<double2>(), n_terms, xm, zm, ny, cre, cim);
    HIPCHECK(hipMemcpyAsync(sv->adj_psi.p, sv->d.p, bytes, hipMemcpyDeviceToDevice, stream()));
    // reverse sweep: gradient point, then the inverse, gate by gate from the end
    reverse_sweep(sv, sv->adj_psi.as<double2>(), sv->adj_lam.as<double2>(), gs, 0, gs.size(), gacc);
  }
  HIPCHECK(hipStreamSynchronize(stream()));
  for (uint64_t k = 0; k < n_params; ++k) grads[k] = gacc[k].value() / 2.0;
  API_END
}

// Adjoint with compressed forward checkpoints: the forward pass stores the
// state at the start of every segment of `segment_gates` gates in a
// compressed store; the reverse sweep restarts each segment from its
// checkpoint (decode, forward over the segment) instead of carrying the
// uncomputed state through the whole circuit, so codec or rounding drift of
// the uncomputed state never crosses a segment boundary.
struct qsim_adj {
  uint32_t n = 0;
  uint64_t seg = 0;
  std::vector<qsim::HGate> gs;
  // checkpoint stores, kept across forward calls while the qubit count,
  // block size and codec settings stay the same (a load resets a store
  // completely), so a repeated forward allocates nothing
  std::vector<qsim_cs*> ckpt;
  uint32_t ck_B = 0;
  qsim_cs_config ck_codec{};
  bool ready = false;
  void clear() {
    for (auto* c : ckpt) qsim_cs_destroy(c);
    ckpt.clear();
    ready = false;
  }
  ~qsim_adj() { clear(); }
};

static bool same_codec(const qsim_cs_config& x, const qsim_cs_config& y) {
  return x.eps_rel == y.eps_rel && x.eps_abs == y.eps_abs && x.guard_bits == y.guard_bits &&
         x.zero_mode == y.zero_mode && x.zero_threshold == y.zero_threshold && x.scratch_bytes == y.scratch_bytes &&
         x.tile_bits == y.tile_bits && x.store_mode == y.store_mode && x.hbm_budget == y.hbm_budget;
}

int qsim_adj_create(qsim_adj** out) {
  API_BEGIN
  require(out != nullptr, "null argument");
  *out = new qsim_adj();
  API_END
}

int qsim_adj_destroy(qsim_adj* adj) {
  API_BEGIN
  delete adj;
  API_END
}

int qsim_adj_forward(qsim_adj* adj, qsim_sv* sv, const qsim_gate* gates, uint64_t n_gates, const double* pool,
                     uint64_t pool_len, uint64_t segment_gates, uint32_t block_size, const qsim_cs_config* codec) {
  API_BEGIN
  require(adj != nullptr && sv != nullptr && codec != nullptr, "null argument");
  require(segment_gates >= 1, "segment_gates must be at least 1");
  require(sv->n >= 5, "checkpointed adjoint needs at least 5 qubits");
  std::vector<HGate> gs = parse_gates(gates, n_gates, pool, pool_len);
  check_qubits(gs, sv->n);
  const uint32_t B = std::min<uint32_t>(std::max<uint32_t>(block_size, 5), sv->n);
  if (adj->n != sv->n || adj->ck_B != B || !same_codec(adj->ck_codec, *codec)) adj->clear();
  adj->ready = false;
  adj->gs = std::move(gs);
  adj->n = sv->n;
  adj->seg = segment_gates;
  adj->ck_B = B;
  adj->ck_codec = *codec;
  const uint64_t G = adj->gs.size();
  size_t k = 0;
  for (uint64_t b = 0; b < G || b == 0; b += segment_gates, ++k) {
    if (k == adj->ckpt.size()) {
      qsim_cs* c = nullptr;
      if (qsim_cs_create(sv->n, B, codec, &c) != 0) throw std::runtime_error(g_err);
      adj->ckpt.push_back(c);
    }
    cs_load_device(adj->ckpt[k], sv->ptr());
    apply_gates_raw(sv->ptr(), sv->n, adj->gs, b, std::min<uint64_t>(G, b + segment_gates));
    if (G == 0) {
      ++k;
      break;
    }
  }
  // stores beyond this circuit's checkpoints
  while (adj->ckpt.size() > k) {
    qsim_cs_destroy(adj->ckpt.back());
    adj->ckpt.pop_back();
  }
  adj->ready = true;
  API_END
}

int qsim_adj_backward(qsim_adj* adj, qsim_sv* sv, uint64_t n_terms, const uint64_t* xm, const uint64_t* zm,
                      const uint32_t* ny, const double* cre, const double* cim, uint64_t n_params, double* energy,
                      double* grads) {
  API_BEGIN
  require(adj != nullptr && sv != nullptr && energy != nullptr && (grads != nullptr || n_params == 0),
          "null argument");
  require(adj->ready && adj->n == sv->n, "qsim_adj_backward needs a qsim_adj_forward on this state first");
  validate_backprop_gates(adj->gs, n_params);
  const uint64_t bytes = sv->dim * sizeof(double2);
  sv->adj_lam.ensure(bytes);
  sv->adj_psi.ensure(bytes);
  *energy = lambda_and_energy(sv, sv->ptr(), sv->adj_lam.as<double2>(), n_terms, xm, zm, ny, cre, cim);
  std::vector<DDAcc> gacc(n_params);
  const uint64_t G = adj->gs.size();
  for (size_t s = adj->ckpt.size(); s-- > 0;) {
    const uint64_t b = (uint64_t)s * adj->seg, e = std::min<uint64_t>(G, b + adj->seg);
    cs_read_device(adj->ckpt[s], sv->adj_psi.as<double2>());
    apply_gates_raw(sv->adj_psi.as<double2>(), sv->n, adj->gs, b, e);
    reverse_sweep(sv, sv->adj_psi.as<double2>(), sv->adj_lam.as<double2>(), adj->gs, b, e, gacc);
  }
  HIPCHECK(hipStreamSynchronize(stream()));
  for (uint64_t k = 0; k < n_params; ++k) grads[k] = gacc[k].value() / 2.0;
  API_END
}

// bytes held by the checkpoints (payload words and descriptors of each store)
int qsim_adj_checkpoint_bytes(const qsim_adj* adj, uint64_t* n_checkpoints, uint64_t* bytes) {
  API_BEGIN
  require(adj != nullptr, "null argument");
  if (n_checkpoints) *n_checkpoints = adj->ckpt.size();
  uint64_t tot = 0;
  for (auto* c : adj->ckpt) tot += c->cur_bytes;
  if (bytes) *bytes = tot;
  API_END
}

int qsim_sv_apply_pauli_sum(const qsim_sv* sv, qsim_sv* out, uint64_t n_terms, const uint64_t* xm, const uint64_t* zm,
                            const uint32_t* ny, const double* cre, const double* cim) {
  API_BEGIN
  require(sv != out, "apply_pauli_sum: output must differ from input");
  require(sv->n == out->n, "qubit count mismatch");
  if (n_terms == 0) HIPCHECK(launch_fill_u64(out->d.as<unsigned long long>(), out->dim * 2, 0ull, stream()));
  for (uint64_t t = 0; t < n_terms; ++t) {
    double pr, pi;
    ipow(ny[t], pr, pi);
    const double fr = cre[t] * pr - cim[t] * pi, fi = cre[t] * pi + cim[t] * pr;
    HIPCHECK(launch_sv_apply_pauli(sv->ptr(), out->ptr(), sv->dim, xm[t], zm[t], fr, fi, t > 0 ? 1 : 0, stream()));
  }
  HIPCHECK(hipStreamSynchronize(stream()));
  API_END
}

void* qsim_sv_device_ptr(qsim_sv* sv) { return sv ? sv->d.p : nullptr; }
uint32_t qsim_sv_qubit_count(const qsim_sv* sv) { return sv ? sv->n : 0; }

// ---------------------------------------------------------------- CS API
int qsim_cs_default_config(qsim_cs_config* cfg) {
  API_BEGIN
  qsim_cs_config c{};
  c.eps_rel = 1e-3;
  c.eps_abs = 1e-7;
  c.guard_bits = 1;
  c.zero_mode = QSIM_ZERO_EXACT;
  c.zero_threshold = 1e-6;
  c.scratch_bytes = 0;
  c.tile_bits = 0;
  *cfg = c;
  API_END
}

int qsim_cs_create(uint32_t n, uint32_t B, const qsim_cs_config* cfg, qsim_cs** out) {
  API_BEGIN
  require(n >= 5 && n <= 44, "qubit count out of range (5..44)");
  require(B >= (uint32_t)kChunkAmpsLog2 && B <= n, "block_size must satisfy 5 <= block_size <= qubit_count");
  require(n - B <= 30, "too many blocks (qubit_count - block_size > 30)");
  require(B <= 26, "block_size above 26");
  qsim_cs_config c;
  if (cfg) c = *cfg;
  else qsim_cs_default_config(&c);
  // eps_rel = 0 is the pure absolute bound (the CPU path's SZ3 ABS mode,
  // compressed_state.hpp:162-189); a zero total budget has no quantiser
  if (!(c.eps_rel >= 0.0) || !(c.eps_abs >= 0.0) || !(c.eps_rel + c.eps_abs > 0.0))
    throw std::invalid_argument("error bounds must be >= 0 with eps_rel + eps_abs > 0");
  require(c.eps_abs >= 0.0, "eps_abs must be >= 0");
  require(c.guard_bits >= 0 && c.guard_bits <= 20, "guard_bits out of range");
  require(c.zero_mode == QSIM_ZERO_EXACT || c.zero_mode == QSIM_ZERO_THRESHOLD, "bad zero_mode");
  require(c.store_mode == QSIM_STORE_DOUBLE || c.store_mode == QSIM_STORE_REGIONS, "bad store_mode");
  auto cs = std::make_unique<qsim_cs>();
  cs->n = n;
  cs->B = B;
  cs->cfg = c;
  cs->nb = 1ull << (n - B);
  cs->dim = 1ull << n;
  cs->cpb = 1ull << (B - kChunkAmpsLog2);
  const uint64_t raw = cs->dim * 16;
  cs->regions = c.store_mode == QSIM_STORE_REGIONS;
  uint64_t sbytes = c.scratch_bytes;
  if (sbytes == 0) {
    const char* e = std::getenv("QOMPSIM_SCRATCH_GIB");
    sbytes = e ? (uint64_t)(std::atof(e) * (1ull << 30)) : (8ull << 30);
    if (!e && !cs->regions) {
      // The double store's multi-pass stages stream their groups through
      // this buffer in batches: larger batches are fewer, longer launches
      // (QFT-32's five-pass stage: 156 ms at 8 GiB, 143 ms with the whole
      // 64 GiB state in one batch).  Default: a quarter of the free device
      // memory (and of the HBM budget), a power of two, at least 8 GiB.
      size_t fr = 0, tot = 0;
      if (hipMemGetInfo(&fr, &tot) == hipSuccess) {
        uint64_t q = fr / 4;
        if (c.hbm_budget) q = std::min<uint64_t>(q, c.hbm_budget / 4);
        uint64_t p2 = 1;
        while (p2 <= q / 2) p2 <<= 1;
        sbytes = std::max<uint64_t>(sbytes, p2);
      }
    }
  }
  sbytes = std::min<uint64_t>(sbytes, raw);
  sbytes = std::max<uint64_t>(sbytes, (1ull << B) * 16);
  // the fixed part of the budget: two descriptor tables and the staging buffer
  const uint64_t fixed = (cs->regions ? 1 : 2) * cs->nb * cs->cpb * 8 + sbytes;
  if (c.hbm_budget && fixed >= c.hbm_budget)
    throw std::invalid_argument("hbm_budget " + std::to_string(c.hbm_budget) +
                                " is below the descriptor tables and staging buffer (" + std::to_string(fixed) +
                                " bytes)");
  cs->scratch.alloc(sbytes);
  cs->scratch_amps = sbytes / 16;
  if (cs->regions) rs_init(cs.get());
  for (int w = 0; w < 2; ++w) {
    if (w == 0 || !cs->regions) cs->desc[w].alloc(cs->nb * cs->cpb * 8);
    cs->cursors[w].alloc((size_t)kMaxSubArenas * 16 * 8);
  }
  if (!cs->regions) {
    // initial capacity (within the budget); an overflowing stage regrows its
    // output arena to the measured demand (QOMPSIM_ARENA_KIB sets the start,
    // for tests)
    uint64_t init = std::min<uint64_t>(std::max<uint64_t>(raw / 16, 8ull << 20), 4ull << 30);
    if (const char* e = std::getenv("QOMPSIM_ARENA_KIB")) init = std::max<uint64_t>(1, std::atoll(e)) << 10;
    if (c.hbm_budget) init = std::min<uint64_t>(init, (c.hbm_budget - fixed) / 4);
    for (int w = 0; w < 2; ++w) cs_alloc_arena(cs.get(), w, words_for_cap(std::max<uint64_t>(init, 512)));
  }
  cs->null_flags.alloc(cs->nb + 8);  // whole dwords are read around a flag
  cs->nonzero.alloc(cs->nb);
  cs->overflow.alloc(16);
  cs->cstats.alloc(8 * (2 + 2 * 64));  // sum and maximum, then 64 partial pairs
  cs->h_null.assign(cs->nb, 1);
  HIPCHECK(hipEventCreate(&cs->ev0));
  HIPCHECK(hipEventCreate(&cs->ev1));
  // io program: one block, no ops, exact zero detection
  plan_program(cs->io, {}, B, std::min<uint32_t>(c.tile_bits ? c.tile_bits : default_tile_bits(), B), kChunkAmpsLog2,
               true);
  cs_fill_codec_fields(cs.get(), cs->io, 0, QSIM_ZERO_EXACT);
  cs->io.upload_all();
  cs_load(cs.get(), nullptr, 0);
  *out = cs.release();
  API_END
}

int qsim_cs_destroy(qsim_cs* cs) {
  API_BEGIN
  if (cs) {
    if (cs->ev0) (void)hipEventDestroy(cs->ev0);
    if (cs->ev1) (void)hipEventDestroy(cs->ev1);
    for (auto e : cs->pev) (void)hipEventDestroy(e);
  }
  delete cs;
  API_END
}

namespace qsim {
// Commutation-aware stage packing (B-aware reordering; no reference
// counterpart, the role the optimizer's swap insertion plays for ranks,
// circuit_optimizer.cpp:834-943): a stage takes, in circuit order, every
// remaining gate whose qubits at or above B fit the stage's high set (at most
// max(I, 2) qubits, the partition_circuit threshold) and that commutes with
// every earlier gate left for a later stage; the rest wait.  Gates keep their
// circuit order inside a stage, and a gate only overtakes gates it commutes
// with, so the product of the stages equals the circuit.
static std::vector<std::vector<uint32_t>> partition_commuting(const std::vector<HGate>& gs, uint32_t B,
                                                              uint32_t inner_size) {
  const uint32_t thr = std::max<uint32_t>(inner_size, 2u);
  std::vector<std::vector<uint32_t>> hi(gs.size());
  for (size_t i = 0; i < gs.size(); ++i) {
    for (auto q : gs[i].t)
      if (q >= B && std::find(hi[i].begin(), hi[i].end(), q) == hi[i].end()) hi[i].push_back(q);
    for (auto q : gs[i].c)
      if (q >= B && std::find(hi[i].begin(), hi[i].end(), q) == hi[i].end()) hi[i].push_back(q);
    if (hi[i].size() > thr) throw std::invalid_argument("gate touches more high qubits than the stage threshold");
  }
  std::vector<std::vector<uint32_t>> stages;
  std::vector<uint32_t> rem(gs.size());
  std::iota(rem.begin(), rem.end(), 0u);
  while (!rem.empty()) {
    std::vector<uint32_t> H, cur, deferred;
    for (uint32_t i : rem) {
      bool ok = true;
      for (uint32_t d : deferred)
        if (!gates_commute(gs[i], gs[d])) { ok = false; break; }
      if (ok) {
        std::vector<uint32_t> H2 = H;
        for (auto q : hi[i])
          if (std::find(H2.begin(), H2.end(), q) == H2.end()) H2.push_back(q);
        if (H2.size() <= thr) {
          H = H2;
          cur.push_back(i);
          continue;
        }
      }
      deferred.push_back(i);
    }
    if (cur.empty()) throw std::logic_error("stage packing made no progress");
    stages.push_back(cur);
    rem.swap(deferred);
  }
  return stages;
}
}  // namespace qsim

int qsim_partition_commuting(uint32_t n_qubits, uint32_t block_size, uint32_t inner_size, const qsim_gate* gates,
                             uint64_t n_gates, const double* pool, uint64_t pool_len, uint32_t* stage_of_gate,
                             uint32_t* n_stages) {
  API_BEGIN
  require(stage_of_gate != nullptr || n_gates == 0, "null argument");
  require(block_size <= n_qubits, "block_size above qubit_count");
  auto gs = parse_gates(gates, n_gates, pool, pool_len);
  check_qubits(gs, n_qubits);
  auto st = partition_commuting(gs, block_size, inner_size);
  for (uint32_t s = 0; s < st.size(); ++s)
    for (auto g : st[s]) stage_of_gate[g] = s;
  if (n_stages) *n_stages = (uint32_t)st.size();
  API_END
}

int qsim_cs_partition_commuting_and_register(qsim_cs* cs, const qsim_gate* gates, uint64_t n_gates, const double* pool,
                                             uint64_t pool_len, uint32_t inner_size) {
  API_BEGIN
  auto gs = parse_gates(gates, n_gates, pool, pool_len);
  check_qubits(gs, cs->n);
  auto st = partition_commuting(gs, cs->B, inner_size);
  cs->gates = gs;
  cs_register(cs, std::move(gs), st, nullptr, nullptr);
  API_END
}

int qsim_cs_partition_and_register(qsim_cs* cs, const qsim_gate* gates, uint64_t n_gates, const double* pool,
                                   uint64_t pool_len, uint32_t inner_size) {
  API_BEGIN
  auto gs = parse_gates(gates, n_gates, pool, pool_len);
  check_qubits(gs, cs->n);
  cs->gates = gs;
  auto st = partition(cs, inner_size);
  cs_register(cs, std::move(gs), st, nullptr, nullptr);
  API_END
}

int qsim_cs_register_stages(qsim_cs* cs, const qsim_gate* gates, uint64_t n_gates, const double* pool,
                            uint64_t pool_len, uint32_t n_stages, const uint32_t* off, const uint32_t* ids,
                            const uint32_t* b2g, const uint32_t* ng) {
  API_BEGIN
  auto gs = parse_gates(gates, n_gates, pool, pool_len);
  std::vector<std::vector<uint32_t>> st(n_stages), ub(n_stages);
  std::vector<uint32_t> ung(n_stages);
  for (uint32_t s = 0; s < n_stages; ++s) {
    st[s].assign(ids + off[s], ids + off[s + 1]);
    ub[s].assign(b2g + (uint64_t)s * cs->nb, b2g + (uint64_t)(s + 1) * cs->nb);
    ung[s] = ng[s];
  }
  cs_register(cs, std::move(gs), st, &ub, &ung);
  API_END
}

int qsim_cs_num_stages(const qsim_cs* cs, uint32_t* out) {
  API_BEGIN
  *out = (uint32_t)cs->stages.size();
  API_END
}

int qsim_cs_stage_info(const qsim_cs* cs, uint32_t s, uint32_t* n_gates, uint32_t* gate_ids, uint32_t* num_groups,
                       uint32_t* b2g, uint32_t* n_passes) {
  API_BEGIN
  if (s >= cs->stages.size()) throw std::out_of_range("stage index out of range");
  const CStage& S = cs->stages[s];
  if (n_gates) *n_gates = (uint32_t)S.gate_ids.size();
  if (gate_ids) std::copy(S.gate_ids.begin(), S.gate_ids.end(), gate_ids);
  if (num_groups) *num_groups = S.num_groups;
  if (b2g) std::copy(S.b2g.begin(), S.b2g.end(), b2g);
  if (n_passes) *n_passes = (uint32_t)S.prog->passes.size();
  API_END
}

int qsim_cs_run_stages(qsim_cs* cs, uint32_t b, uint32_t e) {
  API_BEGIN
  require(b <= e && e <= cs->stages.size(), "stage range out of bounds");
  for (uint32_t s = b; s < e; ++s) cs_run_stage(cs, s);
  API_END
}

int qsim_cs_apply_circuit(qsim_cs* cs) {
  API_BEGIN
  cs->stats.clear();
  // append_total_SV_bytes before the first stage (compressed_state.hpp:880)
  cs->bytes_hist.push_back(cs->cur_bytes);
  for (uint32_t s = 0; s < cs->stages.size(); ++s) cs_run_stage(cs, s);
  API_END
}

static void cs_norm_inner(qsim_cs* cs, const qsim_sv* sv, DDAcc* acc) {
  const uint64_t bpb = cs_blocks_per_batch(cs);
  HIPCHECK(hipMemcpyAsync(cs->h_null.data(), cs->null_flags.p, cs->nb, hipMemcpyDeviceToHost, stream()));
  HIPCHECK(hipStreamSynchronize(stream()));
  for (uint64_t b0 = 0; b0 < cs->nb; b0 += bpb) {
    const uint64_t cnt = std::min(bpb, cs->nb - b0);
    bool any = false;
    for (uint64_t b = b0; b < b0 + cnt; ++b) any = any || !cs->h_null[b];
    if (!any) continue;  // null blocks decode to zeros: no contribution
    std::vector<uint32_t> ids(cnt);
    std::iota(ids.begin(), ids.end(), (uint32_t)b0);
    cs_decode_to_scratch(cs, ids);
    const uint64_t n = cnt << cs->B;
    const unsigned grid = reduce_grid(n);
    if (sv) {
      cs->part.ensure(2, grid);
      HIPCHECK(launch_sv_inner(sv->ptr() + (b0 << cs->B), cs->scratch.as<double2>(), n, cs->part.d.as<double>(), grid,
                               stream()));
      cs->part.fetch(2, grid);
      cs->part.accumulate(2, grid, acc);
    } else {
      cs->part.ensure(1, grid);
      HIPCHECK(launch_sv_norm(cs->scratch.as<double2>(), n, cs->part.d.as<double>(), grid, stream()));
      cs->part.fetch(1, grid);
      cs->part.accumulate(1, grid, acc);
    }
  }
}

int qsim_cs_expectation_pauli(qsim_cs* cs, uint64_t n_terms, const uint64_t* xm, const uint64_t* zm,
                              const uint32_t* ny, const double* cre, const double* cim, double* ore, double* oim) {
  API_BEGIN
  // batches of a power-of-two number of blocks, so a batch starting at block
  // b0 covers global indices (b0 << B) | local and flips stay inside it
  uint64_t bpb = 1;
  while (bpb * 2 <= cs_blocks_per_batch(cs)) bpb *= 2;
  const uint64_t span = bpb << cs->B;
  for (uint64_t t = 0; t < n_terms; ++t)
    if (xm[t] >= span)
      throw std::invalid_argument("expectation: X/Y factors on qubits above the decode batch are not supported");
  std::vector<DDAcc> acc(2 * n_terms);
  HIPCHECK(hipMemcpyAsync(cs->h_null.data(), cs->null_flags.p, cs->nb, hipMemcpyDeviceToHost, stream()));
  HIPCHECK(hipStreamSynchronize(stream()));
  for (uint64_t b0 = 0; b0 < cs->nb; b0 += bpb) {
    const uint64_t cnt = std::min(bpb, cs->nb - b0);
    bool any = false;
    for (uint64_t b = b0; b < b0 + cnt; ++b) any = any || !cs->h_null[b];
    if (!any) continue;
    std::vector<uint32_t> ids(cnt);
    std::iota(ids.begin(), ids.end(), (uint32_t)b0);
    cs_decode_to_scratch(cs, ids);
    const uint64_t n = cnt << cs->B;
    const unsigned grid = reduce_grid(n);
    cs->part.ensure(2, grid);
    for (uint64_t t = 0; t < n_terms; ++t) {
      HIPCHECK(launch_sv_pauli(cs->scratch.as<double2>(), n, xm[t], zm[t], cs->part.d.as<double>(), grid, stream()));
      cs->part.fetch(2, grid);
      DDAcc a[2];
      cs->part.accumulate(2, grid, a);
      const double sg = (__builtin_popcountll((b0 << cs->B) & zm[t]) & 1) ? -1.0 : 1.0;
      acc[2 * t].add(sg * a[0].value());
      acc[2 * t + 1].add(sg * a[1].value());
    }
  }
  DDAcc tr, ti;
  for (uint64_t t = 0; t < n_terms; ++t) {
    double pr, pi;
    ipow(ny[t], pr, pi);
    const double sr = acc[2 * t].value(), si = acc[2 * t + 1].value();
    const double fr = cre[t] * pr - cim[t] * pi, fi = cre[t] * pi + cim[t] * pr;
    tr.add(fr * sr - fi * si);
    ti.add(fr * si + fi * sr);
  }
  *ore = tr.value();
  *oim = ti.value();
  API_END
}

int qsim_cs_norm_squared(qsim_cs* cs, double* out) {
  API_BEGIN
  DDAcc acc;
  cs_norm_inner(cs, nullptr, &acc);
  *out = acc.value();
  API_END
}

int qsim_cs_inner_product_sv(qsim_cs* cs, const qsim_sv* sv, double* re, double* im) {
  API_BEGIN
  require(sv->n == cs->n, "qubit count mismatch");
  DDAcc acc[2];
  cs_norm_inner(cs, sv, acc);
  *re = acc[0].value();
  *im = acc[1].value();
  API_END
}

int qsim_cs_read_vector(qsim_cs* cs, double* host, int normalise) {
  API_BEGIN
  const uint64_t bpb = cs_blocks_per_batch(cs);
  for (uint64_t b0 = 0; b0 < cs->nb; b0 += bpb) {
    const uint64_t cnt = std::min(bpb, cs->nb - b0);
    std::vector<uint32_t> ids(cnt);
    std::iota(ids.begin(), ids.end(), (uint32_t)b0);
    cs_decode_to_scratch(cs, ids);
    HIPCHECK(hipStreamSynchronize(stream()));
    WallTimer t(cs->timers[T_D2H]);
    HIPCHECK(hipMemcpyAsync(host + 2 * (b0 << cs->B), cs->scratch.p, (cnt << cs->B) * 16, hipMemcpyDeviceToHost,
                            stream()));
    HIPCHECK(hipStreamSynchronize(stream()));
  }
  if (normalise) {
    // normalise as csim normalize (update_ops_named_state.cpp): x *= sqrt(1/norm)
    DDAcc acc;
    for (uint64_t i = 0; i < 2 * cs->dim; ++i) acc.add(host[i] * host[i]);
    const double nrm = acc.value();
    if (nrm > 0.0) {
      const double f = std::sqrt(1.0 / nrm);
      for (uint64_t i = 0; i < 2 * cs->dim; ++i) host[i] *= f;
    }
  }
  API_END
}

int qsim_cs_read_block(qsim_cs* cs, uint64_t block, double* host) {
  API_BEGIN
  if (block >= cs->nb) throw std::out_of_range("block index out of range");
  cs_decode_to_scratch(cs, {(uint32_t)block});
  HIPCHECK(hipStreamSynchronize(stream()));
  WallTimer t(cs->timers[T_D2H]);
  HIPCHECK(hipMemcpyAsync(host, cs->scratch.p, (1ull << cs->B) * 16, hipMemcpyDeviceToHost, stream()));
  HIPCHECK(hipStreamSynchronize(stream()));
  API_END
}

int qsim_cs_set_basis(qsim_cs* cs, uint64_t index) {
  API_BEGIN
  require(index < cs->dim, "basis index out of range");
  cs_load(cs, nullptr, index);
  API_END
}

int qsim_cs_load(qsim_cs* cs, const double* host) {
  API_BEGIN
  cs_load(cs, host, 0);
  API_END
}

// CompressedQuantumState::sampling_blockwise (src/cppsim/compressed_state.hpp:1600-1681),
// single process: block masses from the decoded blocks (compensated on the
// GPU, where the reference adds sequentially), a multinomial split of the
// shots over blocks by sequential binomial draws from mt19937_64(seed), per
// block uniforms from mt19937_64(seed + 1) scaled by the block total and
// searched in the block's cumulative distribution, then one shuffle with the
// first generator.  The draws use the same standard-library distributions as
// the reference, so equal masses and distributions give equal samples.
int qsim_cs_sampling_blockwise(qsim_cs* cs, uint32_t n_shots, uint32_t seed, uint64_t* out) {
  API_BEGIN
  require(cs != nullptr && (out != nullptr || n_shots == 0), "null argument");
  const uint64_t namps = 1ull << cs->B;
  HIPCHECK(hipMemcpyAsync(cs->h_null.data(), cs->null_flags.p, cs->nb, hipMemcpyDeviceToHost, stream()));
  HIPCHECK(hipStreamSynchronize(stream()));
  // block masses
  std::vector<double> mass(cs->nb, 0.0);
  const uint64_t bpb = cs_blocks_per_batch(cs);
  DBuf stage, pairs;
  stage.ensure(bpb * kGradReduceChunks * 16);
  pairs.ensure(bpb * 16);
  std::vector<double> hp(2 * bpb);
  for (uint64_t b0 = 0; b0 < cs->nb; b0 += bpb) {
    const uint64_t cnt = std::min(bpb, cs->nb - b0);
    bool any = false;
    for (uint64_t b = b0; b < b0 + cnt; ++b) any = any || !cs->h_null[b];
    if (!any) continue;
    std::vector<uint32_t> ids(cnt);
    std::iota(ids.begin(), ids.end(), (uint32_t)b0);
    cs_decode_to_scratch(cs, ids);
    HIPCHECK(launch_block_sumsq(cs->scratch.as<double2>(), namps, (uint32_t)cnt, stage.as<double>(), pairs.as<double>(),
                                stream()));
    HIPCHECK(hipMemcpyAsync(hp.data(), pairs.p, cnt * 16, hipMemcpyDeviceToHost, stream()));
    HIPCHECK(hipStreamSynchronize(stream()));
    for (uint64_t i = 0; i < cnt; ++i)
      if (!cs->h_null[b0 + i]) mass[b0 + i] = hp[2 * i] + hp[2 * i + 1];
  }
  double total = 0.0;
  for (double m : mass) total += m;
  if (!(total > 0.0) || !std::isfinite(total))
    throw std::domain_error("CompressedQuantumState::sampling_blockwise: state has no finite positive probability mass");
  // multinomial split of the shots over blocks
  std::mt19937_64 rng(seed);
  std::vector<uint32_t> shots(cs->nb, 0);
  uint32_t left = n_shots;
  double rest = total;
  for (size_t g = 0; g < mass.size() && left > 0; ++g) {
    const double p = rest > 0 ? std::min(1.0, mass[g] / rest) : 0.0;
    std::binomial_distribution<uint32_t> bin(left, p);
    const uint32_t k = (g + 1 == mass.size()) ? left : bin(rng);
    shots[g] = k;
    left -= k;
    rest -= mass[g];
  }
  // per-block draws
  std::mt19937_64 rng_local((uint64_t)seed + 1);
  std::uniform_real_distribution<double> uni(0.0, 1.0);
  const uint32_t nt = block_scan_tiles(namps);
  DBuf cdf, tsum, tincl, du, dout;
  cdf.ensure(namps * 8);
  tsum.ensure((uint64_t)nt * 8);
  tincl.ensure((uint64_t)nt * 8);
  std::vector<uint64_t> mine;
  mine.reserve(n_shots);
  std::vector<double> u;
  for (uint64_t b = 0; b < cs->nb; ++b) {
    const uint32_t k = shots[b];
    if (k == 0) continue;
    u.resize(k);
    for (uint32_t i = 0; i < k; ++i) u[i] = uni(rng_local);
    cs_decode_to_scratch(cs, {(uint32_t)b});
    du.ensure((uint64_t)k * 8);
    dout.ensure((uint64_t)k * 8);
    upload(du, u.data(), (uint64_t)k * 8);
    HIPCHECK(launch_block_sample(cs->scratch.as<double2>(), namps, cdf.as<double>(), tsum.as<double>(),
                                 tincl.as<double>(), du.as<double>(), k, b * namps,
                                 dout.as<unsigned long long>(), stream()));
    const size_t at = mine.size();
    mine.resize(at + k);
    HIPCHECK(hipMemcpyAsync(mine.data() + at, dout.p, (uint64_t)k * 8, hipMemcpyDeviceToHost, stream()));
    HIPCHECK(hipStreamSynchronize(stream()));
  }
  std::shuffle(mine.begin(), mine.end(), rng);
  std::copy(mine.begin(), mine.end(), out);
  API_END
}

int qsim_cs_null_bitmap(const qsim_cs* cs, uint8_t* out) {
  API_BEGIN
  std::copy(cs->h_null.begin(), cs->h_null.end(), out);
  API_END
}

int qsim_cs_block_bytes(qsim_cs* cs, uint64_t* out) {
  API_BEGIN
  cs->bytes_tmp.ensure(cs->nb * 8);
  HIPCHECK(launch_block_bytes(cs_desc(cs, cs->cur).as<uint64_t>(), cs->null_flags.as<uint8_t>(), cs->nb, cs->cpb,
                              cs->bytes_tmp.as<unsigned long long>(), stream()));
  HIPCHECK(hipMemcpyAsync(out, cs->bytes_tmp.p, cs->nb * 8, hipMemcpyDeviceToHost, stream()));
  HIPCHECK(hipStreamSynchronize(stream()));
  API_END
}

// ---- compressed-store checkpoint files (no reference counterpart)
//
// A file holds a 128-byte header (store kind, shape, codec settings, the
// caller's next stage, the compressed size), the store's payload layout (the
// double-buffered store: its sub-arena table and allocation cursors; the
// region store: the current half's region key, span and per-region extents),
// the null flags, the descriptor table and the payload words.  Loading reads
// and validates everything but the payload before it touches the store,
// checks that the file holds exactly the payload the layout announces, and
// only then replaces the store; a failure while the payload is copied resets
// the store to |0...0>.
extern "C++" {
namespace qsim {
struct CkptHeader {
  char magic[8];  // "QSIMCKP2"
  uint32_t version, store, n, B;
  uint64_t nb, cpb;
  double eps_rel, eps_abs, zero_thr;
  int32_t guard, zero_mode;
  uint64_t next_stage, cur_bytes;
  uint64_t aux0;  // double store: sub-arena count; region store: the current half
  uint64_t aux1;  // double store: arena capacity in words; region store: region count
  uint64_t payload_words;
  uint64_t key_bits;  // region store: qubits of the region key
  uint64_t reserved;
};
static_assert(sizeof(CkptHeader) == 128, "checkpoint header layout");

// device <-> file through one host staging buffer of at most 256 MiB
static void dev_to_file(std::FILE* f, const void* dev, uint64_t bytes) {
  const uint64_t chunk = 256ull << 20;
  std::vector<char> h((size_t)std::min<uint64_t>(std::max<uint64_t>(bytes, 1), chunk));
  for (uint64_t o = 0; o < bytes; o += chunk) {
    const uint64_t c = std::min(chunk, bytes - o);
    HIPCHECK(hipMemcpy(h.data(), (const char*)dev + o, c, hipMemcpyDeviceToHost));
    if (std::fwrite(h.data(), 1, c, f) != c) throw std::runtime_error("checkpoint: write failed");
  }
}
static void file_to_dev(std::FILE* f, void* dev, uint64_t bytes) {
  const uint64_t chunk = 256ull << 20;
  std::vector<char> h((size_t)std::min<uint64_t>(std::max<uint64_t>(bytes, 1), chunk));
  for (uint64_t o = 0; o < bytes; o += chunk) {
    const uint64_t c = std::min(chunk, bytes - o);
    if (std::fread(h.data(), 1, c, f) != c) throw std::runtime_error("checkpoint: file is truncated");
    HIPCHECK(hipMemcpy((char*)dev + o, h.data(), c, hipMemcpyHostToDevice));
  }
}
template <class T>
static void host_to_file(std::FILE* f, const T* p, uint64_t count) {
  if (count && std::fwrite(p, sizeof(T), count, f) != count) throw std::runtime_error("checkpoint: write failed");
}
template <class T>
static void file_to_host(std::FILE* f, T* p, uint64_t count) {
  if (count && std::fread(p, sizeof(T), count, f) != count) throw std::runtime_error("checkpoint: file is truncated");
}
struct FileCloser {
  std::FILE* f;
  ~FileCloser() {
    if (f) std::fclose(f);
  }
};
static uint64_t file_size(std::FILE* f) {
  const long pos = std::ftell(f);
  std::fseek(f, 0, SEEK_END);
  const long end = std::ftell(f);
  std::fseek(f, pos, SEEK_SET);
  if (pos < 0 || end < 0) throw std::runtime_error("checkpoint: cannot size the file");
  return (uint64_t)end;
}

// Every chunk of a non-null block must decode inside the payload extents:
// width at most 64 and words [off, off + w) within some extent (the arena,
// or one region's mapped words; every extent keeps 8 words of readable
// padding past its end).
static void ckpt_check_descriptors(const qsim_cs* cs, const std::vector<uint8_t>& nullf,
                                   const std::vector<uint64_t>& desc,
                                   const std::vector<std::pair<uint64_t, uint64_t>>& extents) {
  for (uint64_t b = 0; b < cs->nb; ++b) {
    if (nullf[b] > 1) throw std::invalid_argument("checkpoint: corrupt null flag");
    if (nullf[b]) continue;
    for (uint64_t c = 0; c < cs->cpb; ++c) {
      const uint64_t d = desc[b * cs->cpb + c];
      const uint64_t off = desc_off(d), w = desc_w(d);
      if (w > 64) throw std::invalid_argument("checkpoint: corrupt descriptor (code width)");
      bool ok = false;
      for (const auto& e : extents)
        if (off >= e.first && off + w <= e.second) {
          ok = true;
          break;
        }
      if (!ok) throw std::invalid_argument("checkpoint: corrupt descriptor (payload outside the store)");
    }
  }
}
}  // namespace qsim
}  // extern "C++"

int qsim_cs_save_checkpoint(qsim_cs* cs, const char* path, uint64_t next_stage) {
  API_BEGIN
  require(cs != nullptr && path != nullptr, "null argument");
  HIPCHECK(hipStreamSynchronize(stream()));
  const int c = cs->cur;
  CkptHeader h{};
  std::memcpy(h.magic, "QSIMCKP2", 8);
  h.version = 2;
  h.store = cs->regions ? QSIM_STORE_REGIONS : QSIM_STORE_DOUBLE;
  h.n = cs->n;
  h.B = cs->B;
  h.nb = cs->nb;
  h.cpb = cs->cpb;
  h.eps_rel = cs->cfg.eps_rel;
  h.eps_abs = cs->cfg.eps_abs;
  h.zero_thr = cs->cfg.zero_threshold;
  h.guard = cs->cfg.guard_bits;
  h.zero_mode = cs->cfg.zero_mode;
  h.next_stage = next_stage;
  h.cur_bytes = cs->cur_bytes;
  FileCloser fc{std::fopen(path, "wb")};
  if (!fc.f) throw std::runtime_error(std::string("checkpoint: cannot open ") + path);
  if (!cs->regions) {
    h.aux0 = cs->nsub[c];
    h.aux1 = cs->cap_words[c];
    h.payload_words = cs->cap_words[c] + 8;
    std::vector<uint64_t> sb = cs->h_subbase[c];
    sb.resize((size_t)h.aux0 + 1, 0);
    host_to_file(fc.f, &h, 1);
    host_to_file(fc.f, sb.data(), sb.size());
    dev_to_file(fc.f, cs->cursors[c].p, h.aux0 * 16 * 8);
    dev_to_file(fc.f, cs->null_flags.p, cs->nb);
    dev_to_file(fc.f, cs->desc[c].p, cs->nb * cs->cpb * 8);
    dev_to_file(fc.f, cs->arena[c].p, h.payload_words * 8);
  } else {
    const auto& H = cs->rs.half[c];
    const uint32_t K = (uint32_t)H.used.size();
    h.aux0 = (uint64_t)c;
    h.aux1 = K;
    h.key_bits = H.key.size();
    for (auto u : H.used) h.payload_words += u;
    uint32_t key[kRegionKeyBits] = {};
    for (size_t i = 0; i < H.key.size(); ++i) key[i] = H.key[i];
    const uint64_t span = H.span;
    host_to_file(fc.f, &h, 1);
    host_to_file(fc.f, key, kRegionKeyBits);
    host_to_file(fc.f, &span, 1);
    host_to_file(fc.f, H.used.data(), K);
    host_to_file(fc.f, H.payload.data(), K);
    dev_to_file(fc.f, cs->null_flags.p, cs->nb);
    dev_to_file(fc.f, cs_desc(cs, c).p, cs->nb * cs->cpb * 8);
    for (uint32_t r = 0; r < K; ++r) dev_to_file(fc.f, rs_region_ptr(cs, c, r), H.used[r] * 8);
  }
  if (std::fflush(fc.f) != 0) throw std::runtime_error("checkpoint: write failed");
  API_END
}

int qsim_cs_load_checkpoint(qsim_cs* cs, const char* path, uint64_t* next_stage) {
  API_BEGIN
  require(cs != nullptr && path != nullptr, "null argument");
  FileCloser fc{std::fopen(path, "rb")};
  if (!fc.f) throw std::runtime_error(std::string("checkpoint: cannot open ") + path);
  const uint64_t fsize = file_size(fc.f);
  CkptHeader h{};
  if (fsize < sizeof(h)) throw std::invalid_argument("checkpoint: not a compressed-store checkpoint (magic)");
  file_to_host(fc.f, &h, 1);
  if (std::memcmp(h.magic, "QSIMCKP2", 8) != 0 || h.version != 2)
    throw std::invalid_argument("checkpoint: not a compressed-store checkpoint (magic/version)");
  const uint32_t kind = cs->regions ? QSIM_STORE_REGIONS : QSIM_STORE_DOUBLE;
  if (h.store != kind) throw std::invalid_argument("checkpoint: written by the other payload store kind");
  if (h.n != cs->n || h.B != cs->B || h.nb != cs->nb || h.cpb != cs->cpb)
    throw std::invalid_argument("checkpoint: qubit count or block size differs from this store");
  if (h.eps_rel != cs->cfg.eps_rel || h.eps_abs != cs->cfg.eps_abs || h.guard != cs->cfg.guard_bits ||
      h.zero_mode != cs->cfg.zero_mode || h.zero_thr != cs->cfg.zero_threshold)
    throw std::invalid_argument("checkpoint: codec settings differ from this store");
  // ---- layout, null flags and descriptors into host memory, validated
  std::vector<uint64_t> sb, cur;
  uint32_t key[kRegionKeyBits] = {};
  uint64_t span = 0;
  std::vector<uint64_t> used, payload;
  std::vector<std::pair<uint64_t, uint64_t>> extents;
  uint64_t layout_bytes = 0;
  if (kind == QSIM_STORE_DOUBLE) {
    const uint64_t nsub = h.aux0, cap = h.aux1;
    if (nsub == 0 || nsub > (uint64_t)kMaxSubArenas) throw std::invalid_argument("checkpoint: bad sub-arena count");
    if (cap == 0 || cap >= (1ull << 40) || h.payload_words != cap + 8)
      throw std::invalid_argument("checkpoint: bad arena capacity");
    layout_bytes = (nsub + 1) * 8 + nsub * 16 * 8;
    if (fsize < sizeof(h) + layout_bytes) throw std::runtime_error("checkpoint: file is truncated");
    sb.resize(nsub + 1);
    cur.resize(nsub * 16);
    file_to_host(fc.f, sb.data(), sb.size());
    file_to_host(fc.f, cur.data(), cur.size());
    if (sb[0] != 0 || sb[nsub] > cap) throw std::invalid_argument("checkpoint: sub-arena table exceeds the arena");
    for (uint64_t i = 0; i < nsub; ++i) {
      if (sb[i + 1] < sb[i]) throw std::invalid_argument("checkpoint: sub-arena table is not ascending");
      if (cur[i * 16] > sb[i + 1] - sb[i]) throw std::invalid_argument("checkpoint: allocation cursor past its sub-arena");
    }
    extents.emplace_back(0, cap);
  } else {
    const uint64_t hsel = h.aux0, K = h.aux1;
    const uint32_t kb = (uint32_t)h.key_bits;
    if (hsel > 1 || kb > std::min<uint32_t>(kRegionKeyBits, cs->n - cs->B) || K != (1ull << kb))
      throw std::invalid_argument("checkpoint: bad region layout");
    layout_bytes = kRegionKeyBits * 4 + 8 + 2 * K * 8;
    if (fsize < sizeof(h) + layout_bytes) throw std::runtime_error("checkpoint: file is truncated");
    file_to_host(fc.f, key, kRegionKeyBits);
    file_to_host(fc.f, &span, 1);
    used.resize(K);
    payload.resize(K);
    file_to_host(fc.f, used.data(), K);
    file_to_host(fc.f, payload.data(), K);
    for (uint32_t i = 0; i < kb; ++i)
      if (key[i] < cs->B || key[i] >= cs->n || (i > 0 && key[i] <= key[i - 1]))
        throw std::invalid_argument("checkpoint: bad region key");
    const uint64_t rr = (cs->dim * 16) >> kb;
    if (span != (rr + rr / 8 + 2 * cs->rs.unit - 1) / cs->rs.unit * cs->rs.unit)
      throw std::invalid_argument("checkpoint: region span differs from this store");
    uint64_t tot = 0;
    for (uint64_t r = 0; r < K; ++r) {
      if ((used[r] + 8) * 8 > span || payload[r] > used[r])
        throw std::invalid_argument("checkpoint: region extent exceeds its span");
      tot += used[r];
      const uint64_t w0 = (cs->rs.unit + hsel * cs->rs.half_bytes + r * span) / 8;
      extents.emplace_back(w0, w0 + used[r]);
    }
    if (tot != h.payload_words) throw std::invalid_argument("checkpoint: region extents do not add up");
    extents.emplace_back(0, 0);  // chunks without payload point at offset 0
  }
  if (fsize != sizeof(h) + layout_bytes + cs->nb + cs->nb * cs->cpb * 8 + h.payload_words * 8)
    throw std::runtime_error("checkpoint: file size does not match its layout (truncated or corrupt)");
  std::vector<uint8_t> nullf(cs->nb);
  std::vector<uint64_t> desc(cs->nb * cs->cpb);
  file_to_host(fc.f, nullf.data(), nullf.size());
  file_to_host(fc.f, desc.data(), desc.size());
  ckpt_check_descriptors(cs, nullf, desc, extents);
  // ---- replace the store
  HIPCHECK(hipStreamSynchronize(stream()));
  if (kind == QSIM_STORE_DOUBLE) {
    const int c = cs->cur;
    if (cs->cap_words[c] < h.aux1)
      cs_budget_check(cs, (h.aux1 + 8) * 8, cs->arena[c].bytes, "checkpoint arena");
  }
  try {
    if (kind == QSIM_STORE_DOUBLE) {
      const int c = cs->cur;
      if (cs->cap_words[c] < h.aux1) cs_alloc_arena(cs, c, h.aux1);
      cs->nsub[c] = (uint32_t)h.aux0;
      cs->h_subbase[c] = sb;
      cs_upload_subbase(cs, c);
      HIPCHECK(hipMemcpy(cs->cursors[c].p, cur.data(), cur.size() * 8, hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(cs->null_flags.p, nullf.data(), cs->nb, hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(cs->desc[c].p, desc.data(), desc.size() * 8, hipMemcpyHostToDevice));
      file_to_dev(fc.f, cs->arena[c].p, h.payload_words * 8);
    } else {
      const int hs = (int)h.aux0;
      rs_release_half(cs, 0);
      rs_release_half(cs, 1);
      std::vector<uint32_t> kv(key, key + h.key_bits);
      rs_set_layout(cs, hs, kv);
      auto& H = cs->rs.half[hs];
      for (uint32_t r = 0; r < used.size(); ++r) {
        if (used[r]) rs_map(cs, hs, r, used[r]);
        file_to_dev(fc.f, rs_region_ptr(cs, hs, r), used[r] * 8);
        H.used[r] = used[r];
        H.payload[r] = payload[r];
      }
      cs->cur = hs;
      HIPCHECK(hipMemcpy(cs->null_flags.p, nullf.data(), cs->nb, hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(cs_desc(cs, hs).p, desc.data(), desc.size() * 8, hipMemcpyHostToDevice));
    }
  } catch (const std::exception& e) {
    std::string msg = std::string("checkpoint: load failed (") + e.what() + "); the store was reset to |0...0>";
    cs_load(cs, nullptr, 0);
    throw std::runtime_error(msg);
  }
  cs->h_null = nullf;
  cs->cur_bytes = h.cur_bytes;
  if (next_stage) *next_stage = h.next_stage;
  API_END
}

int qsim_cs_bytes_history(const qsim_cs* cs, uint64_t* out, uint64_t cap, uint64_t* n) {
  API_BEGIN
  *n = cs->bytes_hist.size();
  for (uint64_t i = 0; i < std::min<uint64_t>(cap, cs->bytes_hist.size()); ++i) out[i] = cs->bytes_hist[i];
  API_END
}

int qsim_cs_max_sv_bytes(const qsim_cs* cs, uint64_t* out) {
  API_BEGIN
  uint64_t m = 0;
  for (auto v : cs->bytes_hist) m = std::max(m, v);
  *out = m;
  API_END
}

int qsim_cs_store_bytes(const qsim_cs* cs, uint64_t* out3) {
  API_BEGIN
  if (cs->regions) {
    out3[0] = cs->rs.mapped_bytes;
    out3[1] = cs->rs.peak_mapped_bytes;
  } else {
    out3[0] = out3[1] = cs->arena[0].bytes + cs->arena[1].bytes;
  }
  out3[2] = cs->desc[0].bytes + cs->desc[1].bytes + cs->desc_stage.bytes;
  API_END
}

int qsim_cs_stage_stats(const qsim_cs* cs, double* out5, uint64_t cap, uint64_t* n) {
  API_BEGIN
  *n = cs->stats.size();
  for (uint64_t i = 0; i < std::min<uint64_t>(cap, cs->stats.size()); ++i)
    for (int j = 0; j < 5; ++j) out5[i * 5 + j] = cs->stats[i][j];
  API_END
}

int qsim_cs_times(const qsim_cs* cs, double* out) {
  API_BEGIN
  for (int i = 0; i < QSIM_TIMER_COUNT; ++i) out[i] = cs->timers[i];
  API_END
}

uint32_t qsim_cs_qubit_count(const qsim_cs* cs) { return cs ? cs->n : 0; }
uint32_t qsim_cs_block_size(const qsim_cs* cs) { return cs ? cs->B : 0; }

}  // extern "C"

// ================================================================== BATCH
// Many small circuits, each with its own observable, evaluated together:
// energy <H> and the adjoint gradient of every instance.  Every instance is
// planned once (forward passes, lambda = 2 H psi, reverse passes with
// gradient points); the whole sequence is enqueued on the library stream
// with no host synchronisation inside, captured once into a hipGraph and
// replayed.  Work buffers are shared (instances run one after another inside
// the graph); memory is cleared with fill kernels, not memset nodes.
namespace qsim {
struct BInst {
  uint32_t n = 0;
  uint64_t nparams = 0;
  std::vector<HGate> gates;
  std::vector<PauliGroup> G;
  std::vector<PauliTermZ> T;
  DBuf dG, dT;
  Program fwd, rev;
  std::vector<uint32_t> map_off, map_cnt;  // per reverse pass: offset/count into dmap
  std::vector<int32_t> hmap;
  DBuf dmap, dmap_off, dmap_cnt;
  bool planned = false;
  bool persistent = false;  // run by the persistent batch kernel
};

// Small instances (10..16 qubits, wave passes only) run in the persistent
// batch kernel, one work-group each; QOMPSIM_BATCH_PERSISTENT=0 sends every
// instance through its own launches.
static bool batch_persistent_enabled() {
  const char* e = std::getenv("QOMPSIM_BATCH_PERSISTENT");
  return !(e && e[0] == '0');
}

struct TuneChoice {
  int fwd_engine = -1;  // -1: default choice, ENGINE_WAVE or ENGINE_LDS
  uint32_t fwd_tb = 0;  // LDS tile bits when fwd_engine == ENGINE_LDS
  uint32_t adj_tb = 0;  // 0: default
  int adj_engine = -1;  // -1: default choice, ENGINE_WAVE or ENGINE_LDS
};

static void plan_instance(BInst& I, const TuneChoice& tc) {
  std::vector<int> qmap(I.n);
  std::iota(qmap.begin(), qmap.end(), 0);
  std::vector<LGate> lg;
  for (auto& g : I.gates) lg.push_back(lower_gate(g, qmap));
  I.fwd.reset();
  if (!lg.empty()) {
    const bool wave = tc.fwd_engine == ENGINE_LDS ? false : wave_engine_ok(lg, I.n);
    const uint32_t tb = tc.fwd_tb ? tc.fwd_tb : default_tile_bits();
    plan_program_engine(I.fwd, lg, I.n, std::min(tb, I.n), std::min<uint32_t>(5, I.n), false, false, wave);
    I.fwd.upload_all();
  }
  std::vector<LGate> rg;
  for (size_t i = I.gates.size(); i-- > 0;) {
    if (I.gates[i].parametric()) rg.push_back(grad_point(I.gates[i]));
    rg.push_back(lower_gate(inverse_gate(I.gates[i]), qmap));
  }
  I.rev.reset();
  I.map_off.clear();
  I.map_cnt.clear();
  I.hmap.clear();
  // candidates for the persistent kernel plan their reverse sweep on the
  // single-wave 512-amplitude tiles it runs
  const bool want_p = batch_persistent_enabled() && I.n >= kBatchMinQubits && I.n <= kBatchMaxQubits &&
                      tc.fwd_engine != ENGINE_LDS && tc.adj_engine != ENGINE_LDS &&
                      (tc.adj_tb == 0 || tc.adj_tb == kWaveTileSmall) && !I.fwd.passes.empty();
  // (the autotuner leaves these sizes at the default choice: the persistent
  // kernel's gain is throughput over many instances, which a one-instance
  // timing does not see)
  if (!rg.empty()) {
    const uint32_t tb = tc.adj_tb ? std::min(tc.adj_tb, I.n) : adjoint_tile_bits(I.n);
    const bool wave = tc.adj_engine == ENGINE_LDS ? false : adjoint_wave_ok(rg, I.n);
    plan_reverse(I.rev, rg, I.n, tb, wave, want_p ? kWaveTileSmall : 0u);
    for (auto& a : I.rev.passes) {
      I.map_off.push_back((uint32_t)I.hmap.size());
      uint32_t cnt = 0;
      for (uint32_t o = a.op_begin; o < a.op_end; ++o)
        if (I.rev.ops[o].kind == OP_GRAD) {
          I.hmap.push_back((int32_t)I.rev.ops[o].gidx);
          I.rev.ops[o].gidx = cnt++;
        }
      I.map_cnt.push_back(cnt);
    }
    I.rev.upload_all();
    upload(I.dmap, I.hmap.data(), std::max<size_t>(1, I.hmap.size()) * sizeof(int32_t));
    upload(I.dmap_off, I.map_off.data(), std::max<size_t>(1, I.map_off.size()) * sizeof(uint32_t));
    upload(I.dmap_cnt, I.map_cnt.data(), std::max<size_t>(1, I.map_cnt.size()) * sizeof(uint32_t));
  }
  bool ok = want_p && !rg.empty();
  for (auto& a : I.fwd.passes) ok = ok && a.engine == ENGINE_WAVE && a.t == kWaveTile;
  for (auto& a : I.rev.passes) ok = ok && a.engine == ENGINE_WAVE && a.t == kWaveTileSmall;
  for (auto c : I.map_cnt) ok = ok && c <= kBatchMaxPoints;
  I.persistent = ok;
  upload(I.dG, I.G.data(), std::max<size_t>(1, I.G.size()) * sizeof(PauliGroup));
  upload(I.dT, I.T.data(), std::max<size_t>(1, I.T.size()) * sizeof(PauliTermZ));
  HIPCHECK(hipStreamSynchronize(stream()));
  I.planned = true;
}
}  // namespace qsim

struct qsim_batch {
  std::vector<std::unique_ptr<qsim::BInst>> inst;
  std::map<uint32_t, qsim::TuneChoice> tune;
  uint32_t maxn = 0;
  uint64_t maxparams = 0;
  // Instances run on a few independent lanes (lane 0 on the library stream,
  // the others on streams of default priority forked from it and joined back),
  // each with its own working states, so that small instances overlap.
  struct Lane {
    qsim::DBuf psi, psiw, lam, part, ptout, stage;
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    std::vector<size_t> order;  // instances of this lane
  };
  std::vector<std::unique_ptr<Lane>> lanes;
  hipEvent_t start = nullptr;
  qsim::DBuf acc;
  // instances of the persistent kernel: their job table, psi and lambda of
  // each (side by side), and the kernel's instance counter
  std::vector<size_t> pjobs;
  qsim::DBuf jobs, pstates, counter;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool ready = false;  // planned and buffers sized
  ~qsim_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (size_t l = 0; l < lanes.size(); ++l) {
      if (lanes[l]->done) (void)hipEventDestroy(lanes[l]->done);
      if (l > 0 && lanes[l]->st) (void)hipStreamDestroy(lanes[l]->st);
    }
    if (start) (void)hipEventDestroy(start);
  }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};

namespace qsim {
static uint64_t acc_stride(const qsim_batch* b) { return 1 + b->maxparams; }

// size the shared buffers for the planned instances
static uint32_t batch_lanes() {
  const char* e = std::getenv("QOMPSIM_BATCH_LANES");
  const int v = e ? std::atoi(e) : 4;
  return (uint32_t)std::min(std::max(v, 1), 8);
}

static void batch_prepare(qsim_batch* b) {
  uint64_t maxdim = 1, maxpart = 1, maxpts = 1;
  for (auto& I : b->inst) {
    if (!I->planned) plan_instance(*I, b->tune.count(I->n) ? b->tune[I->n] : TuneChoice{});
    const uint64_t dim = 1ull << I->n;
    maxdim = std::max(maxdim, dim);
    maxpart = std::max<uint64_t>(maxpart, reduce_grid(dim));
    for (size_t p = 0; p < I->rev.passes.size(); ++p) {
      const uint64_t ntiles = reverse_tiles(I->rev, p, I->n);
      maxpart = std::max<uint64_t>(maxpart, (uint64_t)I->map_cnt[p] * ntiles);
      maxpts = std::max<uint64_t>(maxpts, I->map_cnt[p]);
    }
  }
  const uint32_t nl = (uint32_t)std::min<size_t>(batch_lanes(), std::max<size_t>(1, b->inst.size()));
  while (b->lanes.size() < nl) {
    auto L = std::make_unique<qsim_batch::Lane>();
    if (b->lanes.empty()) L->st = stream();
    else HIPCHECK(hipStreamCreateWithFlags(&L->st, hipStreamNonBlocking));
    HIPCHECK(hipEventCreateWithFlags(&L->done, hipEventDisableTiming));
    b->lanes.push_back(std::move(L));
  }
  if (!b->start) HIPCHECK(hipEventCreateWithFlags(&b->start, hipEventDisableTiming));
  // largest instances first, each to the lane with the least work so far;
  // the persistent kernel's instances go to its job table (it runs on lane
  // 0, which then takes other instances only when it is the only lane)
  std::vector<size_t> idx(b->inst.size());
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) { return b->inst[x]->n > b->inst[y]->n; });
  b->pjobs.clear();
  for (size_t i : idx)
    if (b->inst[i]->persistent) b->pjobs.push_back(i);
  const size_t lane0 = (!b->pjobs.empty() && nl > 1) ? 1 : 0;
  std::vector<double> work(nl, 0.0);
  for (auto& L : b->lanes) L->order.clear();
  for (size_t i : idx) {
    if (b->inst[i]->persistent) continue;
    const size_t l = std::min_element(work.begin() + lane0, work.begin() + nl) - work.begin();
    b->lanes[l]->order.push_back(i);
    work[l] += std::ldexp(1.0, (int)b->inst[i]->n) + 65536.0;  // amplitudes plus a launch-bound floor
  }
  for (uint32_t l = 0; l < nl; ++l) {
    auto& L = *b->lanes[l];
    L.psi.ensure(maxdim * 16);
    L.psiw.ensure(maxdim * 16);
    L.lam.ensure(maxdim * 16);
    L.part.ensure(maxpart * 16);
    L.ptout.ensure(maxpts * 16);
    L.stage.ensure(std::max<uint64_t>(1, maxpts) * kGradReduceChunks * 16);
  }
  b->acc.ensure(b->inst.size() * acc_stride(b) * 16);
  if (!b->pjobs.empty()) {
    uint64_t words = 0;
    for (size_t i : b->pjobs) words += 2ull << b->inst[i]->n;  // psi and lambda, complex entries
    b->pstates.ensure(words * 16);
    std::vector<BatchJob> jt;
    double2* st = b->pstates.as<double2>();
    for (size_t i : b->pjobs) {
      BInst& I = *b->inst[i];
      BatchJob J{};
      J.n = I.n;
      J.nfwd = (uint32_t)I.fwd.passes.size();
      J.nrev = (uint32_t)I.rev.passes.size();
      J.ng = (uint32_t)I.G.size();
      J.fwd = I.fwd.dargs(0);
      J.rev = I.rev.dargs(0);
      J.groups = I.dG.as<PauliGroup>();
      J.terms = I.dT.as<PauliTermZ>();
      J.map_off = I.dmap_off.as<uint32_t>();
      J.map_cnt = I.dmap_cnt.as<uint32_t>();
      J.pmap = I.dmap.as<int32_t>();
      J.psi = st;
      J.lam = st + (1ull << I.n);
      st += 2ull << I.n;
      J.acc = b->acc.as<double>() + i * acc_stride(b) * 2;
      jt.push_back(J);
    }
    upload(b->jobs, jt.data(), jt.size() * sizeof(BatchJob));
    b->counter.ensure(64);
    if (std::getenv("QOMPSIM_PLAN_LOG"))
      fprintf(stderr, "[batch] persistent kernel instances: %zu of %zu\n", b->pjobs.size(), b->inst.size());
    HIPCHECK(hipStreamSynchronize(stream()));
  }
  b->ready = true;
}

// Enqueue instance i: psi = U|0>, lambda = 2 H psi with <H> into acc[i][0],
// then the reverse sweep accumulating gradients into acc[i][1 + p].
static void batch_enqueue(qsim_batch* b, size_t i, qsim_batch::Lane& L) {
  BInst& I = *b->inst[i];
  const uint64_t dim = 1ull << I.n;
  const hipStream_t st = L.st;
  double2* psi = L.psi.as<double2>();
  double* acc = b->acc.as<double>() + i * acc_stride(b) * 2;
  HIPCHECK(launch_fill_basis(psi, dim, 0, st));
  for (size_t p = 0; p < I.fwd.passes.size(); ++p) {
    PassDyn dy{};
    dy.raw = reinterpret_cast<double*>(psi);
    HIPCHECK(launch_program_pass(I.fwd, p, IN_RAW, OUT_RAW, dy, 1, st));
  }
  const unsigned grid = reduce_grid(dim);
  HIPCHECK(launch_pauli_sum(psi, L.lam.as<double2>(), dim, I.dG.as<PauliGroup>(), (uint32_t)I.G.size(),
                            I.dT.as<PauliTermZ>(), L.part.as<double>(), grid, st));
  HIPCHECK(launch_grad_reduce(L.part.as<double>(), grid, 1, acc, nullptr, st));
  HIPCHECK(hipMemcpyAsync(L.psiw.p, psi, dim * 16, hipMemcpyDeviceToDevice, st));
  for (size_t p = 0; p < I.rev.passes.size(); ++p) {
    const uint64_t ntiles = reverse_tiles(I.rev, p, I.n);
    HIPCHECK(launch_reverse_pass(I.rev, p, I.n, L.psiw.as<double2>(), L.lam.as<double2>(), L.part.as<double>(), st));
    const uint32_t np = I.map_cnt[p];
    if (np == 0) continue;
    HIPCHECK(launch_grad_reduce(L.part.as<double>(), ntiles, np, L.ptout.as<double>(), L.stage.as<double>(), st));
    HIPCHECK(launch_accumulate_points(L.ptout.as<double>(), np, I.dmap.as<int32_t>() + I.map_off[p], acc + 2, st));
  }
}

// The whole batch on the library stream: clear the accumulators, fork the
// other lanes off it, run every lane's instances, join the lanes back.
static void batch_enqueue_all(qsim_batch* b) {
  HIPCHECK(launch_fill_u64(b->acc.as<unsigned long long>(), b->inst.size() * acc_stride(b) * 2, 0ull, stream()));
  HIPCHECK(hipEventRecord(b->start, stream()));
  for (size_t l = 1; l < b->lanes.size(); ++l) HIPCHECK(hipStreamWaitEvent(b->lanes[l]->st, b->start, 0));
  if (!b->pjobs.empty()) {
    // two work-groups per CU at most (256 VGPRs a wave), no more than jobs
    int dev = 0, cus = 0;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const unsigned groups = (unsigned)std::min<uint64_t>(b->pjobs.size(), 2ull * (uint64_t)std::max(cus, 1));
    HIPCHECK(launch_fill_u8(b->counter.as<uint8_t>(), sizeof(unsigned), 0, stream()));
    HIPCHECK(launch_batch_persistent(b->jobs.as<BatchJob>(), (uint32_t)b->pjobs.size(), b->counter.as<unsigned>(),
                                     groups, stream()));
  }
  for (auto& L : b->lanes)
    for (size_t i : L->order) batch_enqueue(b, i, *L);
  for (size_t l = 1; l < b->lanes.size(); ++l) {
    HIPCHECK(hipEventRecord(b->lanes[l]->done, b->lanes[l]->st));
    HIPCHECK(hipStreamWaitEvent(stream(), b->lanes[l]->done, 0));
  }
}

static void batch_read(qsim_batch* b, double* energies, double* grads) {
  std::vector<double> h(b->inst.size() * acc_stride(b) * 2);
  HIPCHECK(hipMemcpyAsync(h.data(), b->acc.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream()));
  HIPCHECK(hipStreamSynchronize(stream()));
  const uint64_t st = acc_stride(b);
  for (size_t i = 0; i < b->inst.size(); ++i) {
    const double* a = h.data() + i * st * 2;
    if (energies) energies[i] = (a[0] + a[1]) / 2.0;
    if (grads)
      for (uint64_t p = 0; p < b->maxparams; ++p)
        grads[i * b->maxparams + p] = p < b->inst[i]->nparams ? (a[2 + 2 * p] + a[3 + 2 * p]) / 2.0 : 0.0;
  }
}
}  // namespace qsim

extern "C" {

int qsim_batch_create(qsim_batch** out) {
  API_BEGIN
  require(out != nullptr, "null argument");
  *out = new qsim_batch();
  API_END
}

int qsim_batch_destroy(qsim_batch* b) {
  API_BEGIN
  if (b) (void)hipStreamSynchronize(stream());
  delete b;
  API_END
}

int qsim_batch_add(qsim_batch* b, uint32_t n_qubits, const qsim_gate* gates, uint64_t n_gates, const double* pool,
                   uint64_t pool_len, uint64_t n_terms, const uint64_t* x_masks, const uint64_t* z_masks,
                   const uint32_t* y_counts, const double* coef_re, const double* coef_im, uint64_t n_params,
                   uint64_t* index) {
  API_BEGIN
  require(b != nullptr, "null batch");
  require(n_qubits >= 4 && n_qubits <= 34, "batch instances need 4..34 qubits");
  auto I = std::make_unique<BInst>();
  I->n = n_qubits;
  I->nparams = n_params;
  I->gates = parse_gates(gates, n_gates, pool, pool_len);
  check_qubits(I->gates, n_qubits);
  for (size_t i = 0; i < I->gates.size(); ++i) {
    const HGate& g = I->gates[i];
    if (!g.parametric()) continue;
    if (g.param < 0 || (uint64_t)g.param >= n_params)
      throw std::out_of_range("gate " + std::to_string(i) + ": parameter index out of range");
    bool ok = g.c.empty() && !g.pauli.empty();
    for (auto pid : g.pauli) ok = ok && pid >= 1 && pid <= 3;
    if (!ok) throw std::invalid_argument("gate " + std::to_string(i) + " does not support backprop (not a Pauli rotation)");
  }
  build_pauli_groups(n_terms, x_masks, z_masks, y_counts, coef_re, coef_im, 2.0, I->G, I->T);
  b->maxn = std::max(b->maxn, n_qubits);
  b->maxparams = std::max(b->maxparams, n_params);
  b->inst.push_back(std::move(I));
  b->ready = false;
  b->drop_graph();
  if (index) *index = b->inst.size() - 1;
  API_END
}

// Autotuner: for every qubit count in the batch, time the forward engine /
// tile choices and the adjoint tile sizes on that count's first instance and
// keep the fastest; later runs plan every instance of that count with it.
int qsim_batch_tune(qsim_batch* b, int reps) {
  API_BEGIN
  require(b != nullptr, "null batch");
  reps = std::max(1, reps);
  std::map<uint32_t, size_t> first;
  for (size_t i = 0; i < b->inst.size(); ++i)
    if (!first.count(b->inst[i]->n)) first[b->inst[i]->n] = i;
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  for (auto& kv : first) {
    const uint32_t n = kv.first;
    BInst& I = *b->inst[kv.second];
    if (batch_persistent_enabled() && n >= kBatchMinQubits && n <= kBatchMaxQubits) {
      TuneChoice p{};
      p.fwd_engine = ENGINE_WAVE;
      p.adj_engine = ENGINE_WAVE;
      p.adj_tb = kWaveTileSmall;
      b->tune[n] = p;
      continue;
    }
    std::vector<TuneChoice> fwd_c, adj_c;
    {
      TuneChoice w;
      w.fwd_engine = ENGINE_WAVE;
      fwd_c.push_back(w);
      for (uint32_t tb : {10u, 12u}) {
        TuneChoice l;
        l.fwd_engine = ENGINE_LDS;
        l.fwd_tb = tb;
        fwd_c.push_back(l);
      }
      for (uint32_t tb : {8u, 9u, 10u, 11u, 12u})
        if (tb <= n) {
          TuneChoice a;
          a.adj_tb = tb;
          a.adj_engine = ENGINE_LDS;
          adj_c.push_back(a);
        }
      if (n >= kWaveTile) {
        TuneChoice a;
        a.adj_tb = kAdjWaveTile;
        a.adj_engine = ENGINE_WAVE;
        adj_c.push_back(a);
      }
    }
    auto time_one = [&](const TuneChoice& tc) {
      plan_instance(I, tc);
      qsim_batch tmp;  // single-instance buffers; the instance itself is borrowed
      tmp.maxparams = std::max<uint64_t>(1, I.nparams);
      tmp.inst.emplace_back(&I);
      float best = 1e30f;
      try {
        batch_prepare(&tmp);
        batch_enqueue_all(&tmp);  // warm: kernel attributes, caches
        for (int r = 0; r < reps; ++r) {
          HIPCHECK(hipEventRecord(e0, stream()));
          batch_enqueue_all(&tmp);
          HIPCHECK(hipEventRecord(e1, stream()));
          HIPCHECK(hipEventSynchronize(e1));
          float ms = 0.f;
          HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
          best = std::min(best, ms);
        }
      } catch (...) {
        (void)tmp.inst.back().release();
        throw;
      }
      (void)tmp.inst.back().release();
      return best;
    };
    TuneChoice best{};
    float bt = 1e30f;
    for (auto& c : fwd_c) {
      const float t = time_one(c);
      if (t < bt) { bt = t; best.fwd_engine = c.fwd_engine; best.fwd_tb = c.fwd_tb; }
    }
    bt = 1e30f;
    TuneChoice fbest = best;
    for (auto& c : adj_c) {
      TuneChoice tc = fbest;
      tc.adj_tb = c.adj_tb;
      tc.adj_engine = c.adj_engine;
      const float t = time_one(tc);
      if (t < bt) { bt = t; best.adj_tb = c.adj_tb; best.adj_engine = c.adj_engine; }
    }
    b->tune[n] = best;
  }
  HIPCHECK(hipEventDestroy(e0));
  HIPCHECK(hipEventDestroy(e1));
  for (auto& I : b->inst) I->planned = false;
  b->ready = false;
  b->drop_graph();
  API_END
}

int qsim_batch_config(const qsim_batch* b, uint32_t n_qubits, int32_t* fwd_engine, uint32_t* fwd_tile_bits,
                      int32_t* adj_engine, uint32_t* adj_tile_bits) {
  API_BEGIN
  require(b != nullptr, "null batch");
  auto it = b->tune.find(n_qubits);
  const TuneChoice tc = it == b->tune.end() ? TuneChoice{} : it->second;
  if (fwd_engine) *fwd_engine = tc.fwd_engine;
  if (fwd_tile_bits) *fwd_tile_bits = tc.fwd_tb;
  if (adj_engine) *adj_engine = tc.adj_engine;
  if (adj_tile_bits) *adj_tile_bits = tc.adj_tb;
  API_END
}

// energies[i] = <H_i>, grads[i * max_params + p] = d<H_i>/d theta_p.  With
// use_graph the first call runs eagerly once, then captures the whole batch
// into a hipGraph; later calls replay it.
int qsim_batch_run(qsim_batch* b, int use_graph, double* energies, double* grads) {
  API_BEGIN
  require(b != nullptr, "null batch");
  if (b->inst.empty()) return 0;
  if (!b->ready) {
    batch_prepare(b);
    b->drop_graph();
  }
  if (!use_graph) {
    batch_enqueue_all(b);
  } else {
    if (!b->exec) {
      batch_enqueue_all(b);  // eager run: validates launches, sets kernel attributes
      HIPCHECK(hipStreamSynchronize(stream()));
      HIPCHECK(hipStreamBeginCapture(stream(), hipStreamCaptureModeRelaxed));
      try {
        batch_enqueue_all(b);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(stream(), &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      HIPCHECK(hipStreamEndCapture(stream(), &b->graph));
      HIPCHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    HIPCHECK(hipGraphLaunch(b->exec, stream()));
  }
  batch_read(b, energies, grads);
  API_END
}

int qsim_batch_traffic(qsim_batch* b, uint64_t* bytes, uint64_t* persistent) {
  API_BEGIN
  require(b != nullptr && bytes != nullptr, "null argument");
  if (!b->ready && !b->inst.empty()) {
    batch_prepare(b);
    b->drop_graph();
  }
  uint64_t t = 0;
  for (auto& I : b->inst) {
    const uint64_t dim = 1ull << I->n;
    t += 16 * dim + 32 * dim * (I->fwd.passes.size() + 1) + 64 * dim * I->rev.passes.size();
  }
  *bytes = t;
  if (persistent) *persistent = b->pjobs.size();
  API_END
}

uint64_t qsim_batch_size(const qsim_batch* b) { return b ? b->inst.size() : 0; }
uint64_t qsim_batch_max_params(const qsim_batch* b) { return b ? b->maxparams : 0; }

}  // extern "C"
