// ORACLE — test infrastructure, not product code.
//
// CPU restatement of the reference's algorithms for the compressed stage loop,
// used only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// as the checker and the timed CPU baseline.  The product (libqsim_hip.so)
// never links, loads or calls it.  Built with g++ -O3 -fopenmp (oracle/build.py).
//
// Parity pins (tests/test_oracle.py): closed forms (GHZ, basis permutations,
// QFT = unitary DFT checked by numpy FFT), kron-built unitaries for small n,
// the reference's own src/csim kernels compiled in place into oracle/_ref by
// oracle/ref_csim.mk when the reference tree is present, the derived GHZ-20
// stage/bitmap fixtures, and codec contract properties.
#include <omp.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <set>
#include <random>
#include <vector>

namespace {
using cplx = std::complex<double>;

struct GateRec {  // same layout as qsim_gate in include/qsim.h
  uint32_t n_targets, n_controls;
  uint32_t targets[8], controls[8], control_values[8], target_commute[8];
  uint32_t flags;
  int32_t param_index;
  uint32_t pauli_ids[8];
  uint64_t matrix_offset;
};
static_assert(sizeof(GateRec) == 184, "qsim_gate layout");

// ----------------------------------------------------------------- kernels
// Dense k-target gate with controls, Qulacs semantics: matrix index bit i <->
// targets[i], applied where every control has its value
// (src/cppsim/gate_matrix.cpp:71-250; src/csim/update_ops_matrix_dense_single.cpp:22-68,
// update_ops_matrix_dense_double.cpp:42-113, update_ops_matrix_dense_multi.cpp:29-48).
void apply_dense(cplx* psi, uint32_t n, const GateRec& g, const cplx* mat) {
  const uint32_t k = g.n_targets, nc = g.n_controls;
  std::vector<uint32_t> all;
  for (uint32_t i = 0; i < k; ++i) all.push_back(g.targets[i]);
  for (uint32_t i = 0; i < nc; ++i) all.push_back(g.controls[i]);
  std::sort(all.begin(), all.end());
  uint64_t cbits = 0;
  for (uint32_t i = 0; i < nc; ++i)
    if (g.control_values[i]) cbits |= 1ull << g.controls[i];
  const uint64_t D = 1ull << k;
  std::vector<uint64_t> off(D);
  for (uint64_t c = 0; c < D; ++c) {
    uint64_t o = 0;
    for (uint32_t j = 0; j < k; ++j)
      if ((c >> j) & 1) o |= 1ull << g.targets[j];
    off[c] = o;
  }
  const uint64_t outer = 1ull << (n - all.size());
#pragma omp parallel if (outer >= 4096)
  {
    std::vector<cplx> v(D), w(D);
#pragma omp for schedule(static)
    for (int64_t oi = 0; oi < (int64_t)outer; ++oi) {
      uint64_t b = (uint64_t)oi;
      for (uint32_t q : all) b = (b & ((1ull << q) - 1)) | ((b >> q) << (q + 1));
      b |= cbits;
      for (uint64_t c = 0; c < D; ++c) v[c] = psi[b | off[c]];
      for (uint64_t r = 0; r < D; ++r) {
        cplx s(0, 0);
        for (uint64_t c = 0; c < D; ++c) s += mat[r * D + c] * v[c];
        w[r] = s;
      }
      for (uint64_t r = 0; r < D; ++r) psi[b | off[r]] = w[r];
    }
  }
}

struct DD {
  double s = 0, c = 0;
  void add(double x) {
    double t = s + x, bp = t - s;
    c += (s - (t - bp)) + (x - bp);
    s = t;
  }
  void merge(const DD& o) { add(o.s); c += o.c; }
  double v() const { return s + c; }
};

// Compensated sums (state_norm_squared, state_inner_product: src/csim/stat_ops.cpp:13-86).
template <class F>
void dd_reduce(uint64_t n, int nv, F f, double* out) {
  int nt = omp_get_max_threads();
  std::vector<DD> acc((size_t)nt * nv);
#pragma omp parallel
  {
    int t = omp_get_thread_num();
    DD* a = &acc[(size_t)t * nv];
#pragma omp for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) f((uint64_t)i, a);
  }
  for (int v = 0; v < nv; ++v) {
    DD tot;
    for (int t = 0; t < nt; ++t) tot.merge(acc[(size_t)t * nv + v]);
    out[v] = tot.v();
  }
}

// ------------------------------------------------------------------- codec
// The error-bounded contract of src/cppsim/compand_bitcomp_pipeline.cuh:14-62
// (|x^ - x| <= eps_rel|x| + eps_abs) met by per-chunk linear quantisation with
// a power-of-two step; one chunk = 64 doubles.
struct Chunk {
  uint32_t w;
  int e;
};

Chunk chunk_plan(const double* x, double er, double ea, int guard, uint64_t* codes) {
  bool any = false;
  for (int l = 0; l < 64; ++l) any |= x[l] != 0.0;
  if (!any) {
    for (int l = 0; l < 64; ++l) codes[l] = 0;
    return {0, 0};
  }
  int emin = INT_MAX;
  bool raw = false;
  for (int l = 0; l < 64; ++l) {
    if (x[l] == 0.0) continue;
    if (!std::isfinite(x[l])) { raw = true; break; }
    const double y = 2.0 * std::fma(er, std::fabs(x[l]), ea);
    if (!(y > 0.0) || !std::isfinite(y)) { raw = true; break; }
    emin = std::min(emin, std::ilogb(y));
  }
  if (raw || emin < -1000) {
    for (int l = 0; l < 64; ++l) std::memcpy(&codes[l], &x[l], 8);
    return {64, 0};
  }
  const int e = emin - guard;
  int bl = 0;
  for (int l = 0; l < 64; ++l) {
    const double s = std::ldexp(x[l], -e);
    if (!(std::fabs(s) < 2251799813685248.0)) {
      for (int m = 0; m < 64; ++m) std::memcpy(&codes[m], &x[m], 8);
      return {64, 0};
    }
    const double k = std::rint(s);
    const double uz = k >= 0.0 ? 2.0 * k : -2.0 * k - 1.0;
    codes[l] = (uint64_t)uz;
    if (codes[l]) bl = std::max(bl, 64 - __builtin_clzll(codes[l]));
  }
  return {(uint32_t)bl, e};
}

double chunk_value(uint64_t u, Chunk c) {
  if (c.w == 0) return 0.0;
  if (c.w == 64) {
    double d;
    std::memcpy(&d, &u, 8);
    return d;
  }
  const int64_t k = (int64_t)(u >> 1) ^ -(int64_t)(u & 1ull);
  return std::ldexp((double)k, c.e);
}

void pack(const uint64_t* codes, uint32_t w, uint64_t* out) {
  for (uint32_t q = 0; q < w; ++q) out[q] = 0;
  for (uint32_t l = 0; l < 64; ++l) {
    const uint32_t bit = l * w, q = bit >> 6, s = bit & 63;
    out[q] |= codes[l] << s;
    if (s + w > 64) out[q + 1] |= codes[l] >> (64 - s);
  }
}

void unpack(const uint64_t* in, uint32_t w, uint64_t* codes) {
  for (uint32_t l = 0; l < 64; ++l) {
    if (w == 64) { codes[l] = in[l]; continue; }
    const uint32_t bit = l * w, q = bit >> 6, s = bit & 63;
    uint64_t v = in[q] >> s;
    if (s + w > 64) v |= in[q + 1] << (64 - s);
    codes[l] = v & ((1ull << w) - 1);
  }
}

struct Block {
  bool null = true;
  std::vector<uint8_t> w;
  std::vector<int16_t> e;
  std::vector<uint32_t> off;   // payload word offset per chunk
  std::vector<uint64_t> payload;
  uint64_t bytes() const { return null ? 0 : 8 * (payload.size() + w.size()); }
};

struct Codec {
  double er, ea;
  int guard;
};

void encode_block(const cplx* amps, uint64_t namps, const Codec& cd, Block& b) {
  const uint64_t nch = namps / 32;
  b.null = false;
  b.w.assign(nch, 0);
  b.e.assign(nch, 0);
  b.off.assign(nch, 0);
  std::vector<Chunk> ch(nch);
  std::vector<uint64_t> codes(nch * 64);
  const double* x = reinterpret_cast<const double*>(amps);
#pragma omp parallel for schedule(static) if (nch >= 64)
  for (int64_t c = 0; c < (int64_t)nch; ++c) ch[c] = chunk_plan(x + 64 * c, cd.er, cd.ea, cd.guard, &codes[64 * c]);
  uint64_t tot = 0;
  for (uint64_t c = 0; c < nch; ++c) {
    b.w[c] = (uint8_t)ch[c].w;
    b.e[c] = (int16_t)ch[c].e;
    b.off[c] = (uint32_t)tot;
    tot += ch[c].w;
  }
  b.payload.assign(tot, 0);
#pragma omp parallel for schedule(static) if (nch >= 64)
  for (int64_t c = 0; c < (int64_t)nch; ++c)
    if (ch[c].w) pack(&codes[64 * c], ch[c].w, &b.payload[b.off[c]]);
}

void decode_block(const Block& b, uint64_t namps, cplx* amps) {
  double* x = reinterpret_cast<double*>(amps);
  if (b.null) {
    std::fill(x, x + 2 * namps, 0.0);
    return;
  }
  const uint64_t nch = namps / 32;
#pragma omp parallel for schedule(static) if (nch >= 64)
  for (int64_t c = 0; c < (int64_t)nch; ++c) {
    uint64_t codes[64];
    const Chunk ck{b.w[c], b.e[c]};
    if (ck.w) unpack(&b.payload[b.off[c]], ck.w, codes);
    for (int l = 0; l < 64; ++l) x[64 * c + l] = ck.w ? chunk_value(codes[l], ck) : 0.0;
  }
}

// -------------------------------------------------------- compressed state
struct Gate {
  GateRec r;
  std::vector<cplx> m;
};

struct CState {
  uint32_t n, B;
  uint64_t nb, namps;
  Codec cd;
  int zero_mode;  // 0 exact, 1 threshold
  double thr;
  std::vector<Block> blocks;
  std::vector<Gate> gates;
  std::vector<std::vector<uint32_t>> stages;
  std::vector<std::vector<uint32_t>> b2g;
  std::vector<uint32_t> ngroups;
  std::vector<uint64_t> bytes_hist;
};

uint64_t total_bytes(const CState& s) {
  uint64_t t = 0;
  for (auto& b : s.blocks) t += b.bytes();
  return t;
}

// partition_circuit (src/cppsim/compressed_state.hpp:509-583)
std::vector<std::vector<uint32_t>> partition_circuit(const CState& s, uint32_t inner_size) {
  std::vector<std::vector<uint32_t>> stages;
  std::vector<uint32_t> current, global_indices;
  const uint32_t threshold = std::max(inner_size, 2u);
  for (uint32_t i = 0; i < s.gates.size(); ++i) {
    const GateRec& g = s.gates[i].r;
    std::vector<uint32_t> gt;
    for (uint32_t j = 0; j < g.n_targets; ++j)
      if (g.targets[j] >= s.B) gt.push_back(g.targets[j]);
    for (uint32_t j = 0; j < g.n_controls; ++j)
      if (g.controls[j] >= s.B) gt.push_back(g.controls[j]);
    std::vector<uint32_t> tmp = global_indices;
    for (auto q : gt)
      if (std::find(tmp.begin(), tmp.end(), q) == tmp.end()) tmp.push_back(q);
    if (tmp.size() > threshold) {
      stages.push_back(current);
      current.clear();
      tmp = gt;
    }
    global_indices = tmp;
    current.push_back(i);
  }
  if (!current.empty()) stages.push_back(current);
  return stages;
}

// make_block_groups (src/cppsim/compressed_state.hpp:585-671)
void make_block_groups(CState& s) {
  s.b2g.clear();
  s.ngroups.clear();
  for (auto& st : s.stages) {
    std::set<uint32_t> inner;
    for (uint32_t i = 0; i < s.B; ++i) inner.insert(i);
    for (auto gi : st) {
      const GateRec& g = s.gates[gi].r;
      for (uint32_t j = 0; j < g.n_targets; ++j)
        if (g.targets[j] >= s.B) inner.insert(g.targets[j]);
      for (uint32_t j = 0; j < g.n_controls; ++j)
        if (g.controls[j] >= s.B) inner.insert(g.controls[j]);
    }
    std::vector<uint32_t> outer;
    for (uint32_t q = 0; q < s.n; ++q)
      if (!inner.count(q)) outer.push_back(q);
    std::vector<uint32_t> ids(s.nb);
    for (uint64_t b = 0; b < s.nb; ++b) {
      uint32_t gid = 0;
      for (uint32_t i = 0; i < outer.size(); ++i) gid |= (uint32_t)((b >> (outer[i] - s.B)) & 1u) << i;
      ids[b] = gid;
    }
    s.b2g.push_back(ids);
    s.ngroups.push_back(1u << outer.size());
  }
}

// process_SV_group (src/cppsim/compressed_state.hpp:953-1029), with the
// group's blocks gathered into one local vector: local qubit q<B is qubit q,
// local qubit B+i is the i-th touched high qubit.
void process_group(CState& s, const std::vector<uint32_t>& stage, const std::vector<uint32_t>& blocks,
                   const std::vector<uint32_t>& high) {
  const uint64_t nloc = blocks.size() * s.namps;
  std::vector<cplx> v(nloc);
  std::vector<bool> was_null(blocks.size());
  for (size_t j = 0; j < blocks.size(); ++j) {
    was_null[j] = s.blocks[blocks[j]].null;
    decode_block(s.blocks[blocks[j]], s.namps, &v[j * s.namps]);
  }
  const uint32_t L = s.B + (uint32_t)high.size();
  for (auto gi : stage) {
    GateRec g = s.gates[gi].r;
    auto loc = [&](uint32_t q) -> uint32_t {
      if (q < s.B) return q;
      for (uint32_t i = 0; i < high.size(); ++i)
        if (high[i] == q) return s.B + i;
      return UINT32_MAX;
    };
    for (uint32_t j = 0; j < g.n_targets; ++j) g.targets[j] = loc(g.targets[j]);
    for (uint32_t j = 0; j < g.n_controls; ++j) g.controls[j] = loc(g.controls[j]);
    apply_dense(v.data(), L, g, s.gates[gi].m.data());
  }
  for (size_t j = 0; j < blocks.size(); ++j) {
    const cplx* a = &v[j * s.namps];
    if (was_null[j]) {
      bool zero = true;
      for (uint64_t i = 0; i < s.namps && zero; ++i) {
        if (s.zero_mode == 0) zero = (a[i].real() == 0.0 && a[i].imag() == 0.0);
        else zero = !(std::fabs(a[i].real()) > s.thr || std::fabs(a[i].imag()) > s.thr);
      }
      if (zero) continue;
    }
    encode_block(a, s.namps, s.cd, s.blocks[blocks[j]]);
  }
}

// apply_circuit (src/cppsim/compressed_state.hpp:840-937) with
// build_stage_group_execution_plan (797-837).
void apply_circuit(CState& s, uint32_t sb, uint32_t se) {
  for (uint32_t si = sb; si < se; ++si) {
    const auto& b2g = s.b2g[si];
    std::vector<uint32_t> cnt(s.ngroups[si], 0);
    for (uint64_t b = 0; b < s.nb; ++b)
      if (!s.blocks[b].null) cnt[b2g[b]]++;
    std::vector<std::vector<uint32_t>> members(s.ngroups[si]);
    for (uint64_t b = 0; b < s.nb; ++b) members[b2g[b]].push_back((uint32_t)b);
    std::set<uint32_t> hs;
    for (auto gi : s.stages[si]) {
      const GateRec& g = s.gates[gi].r;
      for (uint32_t j = 0; j < g.n_targets; ++j)
        if (g.targets[j] >= s.B) hs.insert(g.targets[j]);
      for (uint32_t j = 0; j < g.n_controls; ++j)
        if (g.controls[j] >= s.B) hs.insert(g.controls[j]);
    }
    std::vector<uint32_t> high(hs.begin(), hs.end());
    std::vector<uint32_t> active;
    for (uint32_t gid = 0; gid < s.ngroups[si]; ++gid)
      if (cnt[gid] > 0) active.push_back(gid);
    // groups are independent; a group runs on one thread when there are
    // enough groups, otherwise the group's own loops use the threads
    if (active.size() >= (size_t)omp_get_max_threads()) {
#pragma omp parallel for schedule(dynamic, 1)
      for (int64_t a = 0; a < (int64_t)active.size(); ++a) process_group(s, s.stages[si], members[active[a]], high);
    } else {
      for (auto gid : active) process_group(s, s.stages[si], members[gid], high);
    }
    s.bytes_hist.push_back(total_bytes(s));
  }
}

// backprop / backprop_inner_product (src/vqcsim/parametric_circuit.cpp:180-308)
struct Terms {
  uint64_t nt;
  const uint64_t *xm, *zm;
  const uint32_t* ny;
  const double *cr, *ci;
};

cplx ipow(uint32_t ny) {
  switch (ny & 3u) {
    case 0: return {1, 0};
    case 1: return {0, 1};
    case 2: return {-1, 0};
    default: return {0, -1};
  }
}

void apply_pauli_sum(const cplx* psi, cplx* out, uint64_t dim, const Terms& T) {
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < (int64_t)dim; ++j) {
    cplx acc(0, 0);
    for (uint64_t t = 0; t < T.nt; ++t) {
      const uint64_t i = (uint64_t)j ^ T.xm[t];
      const double sg = (__builtin_popcountll(i & T.zm[t]) & 1) ? -1.0 : 1.0;
      acc += cplx(T.cr[t], T.ci[t]) * ipow(T.ny[t]) * sg * psi[i];
    }
    out[j] = acc;
  }
}

cplx expectation(const cplx* psi, uint64_t dim, const Terms& T) {
  DD re, im;
  for (uint64_t t = 0; t < T.nt; ++t) {
    double r[2];
    dd_reduce(dim, 2, [&](uint64_t i, DD* a) {
      const cplx y = psi[i ^ T.xm[t]];
      cplx v = std::conj(y) * psi[i];
      if (__builtin_popcountll(i & T.zm[t]) & 1) v = -v;
      a[0].add(v.real());
      a[1].add(v.imag());
    }, r);
    const cplx f = cplx(T.cr[t], T.ci[t]) * ipow(T.ny[t]) * cplx(r[0], r[1]);
    re.add(f.real());
    im.add(f.imag());
  }
  return {re.v(), im.v()};
}

std::vector<Gate> load_gates(const GateRec* g, uint64_t n, const double* pool) {
  std::vector<Gate> out(n);
  for (uint64_t i = 0; i < n; ++i) {
    out[i].r = g[i];
    const uint64_t D = 1ull << g[i].n_targets;
    out[i].m.resize(D * D);
    for (uint64_t k = 0; k < D * D; ++k)
      out[i].m[k] = cplx(pool[2 * (g[i].matrix_offset + k)], pool[2 * (g[i].matrix_offset + k) + 1]);
  }
  return out;
}

Gate inverse_of(const Gate& g) {
  Gate h = g;
  const uint64_t D = 1ull << g.r.n_targets;
  for (uint64_t r = 0; r < D; ++r)
    for (uint64_t c = 0; c < D; ++c) h.m[r * D + c] = std::conj(g.m[c * D + r]);
  return h;
}

// R_P(theta) = exp(+i theta P/2) on the gate's targets with its Pauli ids
Gate rotation_of(const Gate& g, double theta) {
  Gate h = g;
  const uint32_t k = g.r.n_targets;
  const uint64_t D = 1ull << k;
  std::vector<cplx> P(D * D, cplx(0, 0));
  for (uint64_t c = 0; c < D; ++c) {
    uint64_t r = c;
    cplx ph(1, 0);
    for (uint32_t j = 0; j < k; ++j) {
      const uint32_t p = g.r.pauli_ids[j];
      const uint64_t b = (c >> j) & 1;
      if (p == 1 || p == 2) r ^= 1ull << j;
      if (p == 2) ph *= b ? cplx(0, -1) : cplx(0, 1);
      if (p == 3 && b) ph *= -1.0;
    }
    P[r * D + c] = ph;
  }
  for (uint64_t r = 0; r < D; ++r)
    for (uint64_t c = 0; c < D; ++c)
      h.m[r * D + c] = (r == c ? std::cos(theta / 2) : 0.0) + cplx(0, std::sin(theta / 2)) * P[r * D + c];
  return h;
}

}  // namespace

extern "C" {

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int n) { omp_set_num_threads(n); }

void orc_apply_circuit(double* psi, uint32_t n, const void* gates, uint64_t ng, const double* pool) {
  auto gs = load_gates(static_cast<const GateRec*>(gates), ng, pool);
  for (auto& g : gs) apply_dense(reinterpret_cast<cplx*>(psi), n, g.r, g.m.data());
}

double orc_norm(const double* psi, uint64_t dim) {
  double r;
  dd_reduce(dim, 1, [&](uint64_t i, DD* a) { a[0].add(psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1]); }, &r);
  return r;
}

void orc_inner(const double* a, const double* b, uint64_t dim, double* re, double* im) {
  double r[2];
  const cplx* A = reinterpret_cast<const cplx*>(a);
  const cplx* Bv = reinterpret_cast<const cplx*>(b);
  dd_reduce(dim, 2, [&](uint64_t i, DD* acc) {
    const cplx v = std::conj(A[i]) * Bv[i];
    acc[0].add(v.real());
    acc[1].add(v.imag());
  }, r);
  *re = r[0];
  *im = r[1];
}

void orc_expectation(const double* psi, uint32_t n, uint64_t nt, const uint64_t* xm, const uint64_t* zm,
                     const uint32_t* ny, const double* cr, const double* ci, double* ore, double* oim) {
  const Terms T{nt, xm, zm, ny, cr, ci};
  const cplx e = expectation(reinterpret_cast<const cplx*>(psi), 1ull << n, T);
  *ore = e.real();
  *oim = e.imag();
}

void orc_apply_pauli_sum(const double* psi, double* out, uint32_t n, uint64_t nt, const uint64_t* xm,
                         const uint64_t* zm, const uint32_t* ny, const double* cr, const double* ci) {
  const Terms T{nt, xm, zm, ny, cr, ci};
  apply_pauli_sum(reinterpret_cast<const cplx*>(psi), reinterpret_cast<cplx*>(out), 1ull << n, T);
}

// Codec of n_doubles (multiple of 64) doubles: decoded values and payload words.
void orc_codec_roundtrip(const double* in, uint64_t nd, double er, double ea, int guard, double* out,
                         uint64_t* payload_words, uint8_t* widths) {
  uint64_t tot = 0;
  for (uint64_t c = 0; c < nd / 64; ++c) {
    uint64_t codes[64], packed[64], back[64];
    const Chunk ck = chunk_plan(in + 64 * c, er, ea, guard, codes);
    if (widths) widths[c] = (uint8_t)ck.w;
    tot += ck.w;
    if (ck.w) {
      pack(codes, ck.w, packed);
      unpack(packed, ck.w, back);
    }
    for (int l = 0; l < 64; ++l) out[64 * c + l] = ck.w ? chunk_value(back[l], ck) : 0.0;
  }
  *payload_words = tot;
}

// Blocks hold whole codec chunks (32 amplitudes): 5 <= B <= n, else null.
void* orc_cs_create(uint32_t n, uint32_t B, double er, double ea, int guard, int zero_mode, double thr) {
  if (B < 5 || B > n) return nullptr;
  auto* s = new CState();
  s->n = n;
  s->B = B;
  s->nb = 1ull << (n - B);
  s->namps = 1ull << B;
  s->cd = {er, ea, guard};
  s->zero_mode = zero_mode;
  s->thr = thr;
  s->blocks.resize(s->nb);
  // initialize_all_blocks (compressed_state.hpp:327-352): block 0 = |0..0>
  std::vector<cplx> tmp(s->namps, cplx(0, 0));
  tmp[0] = 1.0;
  encode_block(tmp.data(), s->namps, s->cd, s->blocks[0]);
  return s;
}

void orc_cs_destroy(void* h) { delete static_cast<CState*>(h); }

void orc_cs_load(void* h, const double* vec) {
  CState& s = *static_cast<CState*>(h);
  const cplx* v = reinterpret_cast<const cplx*>(vec);
  for (uint64_t b = 0; b < s.nb; ++b) {
    const cplx* a = v + b * s.namps;
    bool zero = true;
    for (uint64_t i = 0; i < s.namps && zero; ++i) zero = (a[i] == cplx(0, 0));
    if (zero) { s.blocks[b] = Block(); continue; }
    encode_block(a, s.namps, s.cd, s.blocks[b]);
  }
}

uint32_t orc_cs_partition_and_register(void* h, const void* gates, uint64_t ng, const double* pool, uint32_t inner) {
  CState& s = *static_cast<CState*>(h);
  s.gates = load_gates(static_cast<const GateRec*>(gates), ng, pool);
  s.stages = partition_circuit(s, inner);
  make_block_groups(s);
  return (uint32_t)s.stages.size();
}

uint32_t orc_cs_stage(void* h, uint32_t si, uint32_t* gate_ids, uint32_t* b2g, uint32_t* num_groups) {
  CState& s = *static_cast<CState*>(h);
  if (gate_ids) std::copy(s.stages[si].begin(), s.stages[si].end(), gate_ids);
  if (b2g) std::copy(s.b2g[si].begin(), s.b2g[si].end(), b2g);
  if (num_groups) *num_groups = s.ngroups[si];
  return (uint32_t)s.stages[si].size();
}

void orc_cs_apply_circuit(void* h) {
  CState& s = *static_cast<CState*>(h);
  s.bytes_hist.push_back(total_bytes(s));
  apply_circuit(s, 0, (uint32_t)s.stages.size());
}

void orc_cs_run_stages(void* h, uint32_t b, uint32_t e) { apply_circuit(*static_cast<CState*>(h), b, e); }

void orc_cs_read_vector(void* h, double* out, int normalise) {
  CState& s = *static_cast<CState*>(h);
  cplx* o = reinterpret_cast<cplx*>(out);
  for (uint64_t b = 0; b < s.nb; ++b) decode_block(s.blocks[b], s.namps, o + b * s.namps);
  if (normalise) {
    const double nrm = orc_norm(out, 1ull << s.n);
    const double f = std::sqrt(1.0 / nrm);
    for (uint64_t i = 0; i < 2 * (1ull << s.n); ++i) out[i] *= f;
  }
}

void orc_cs_null_bitmap(void* h, uint8_t* out) {
  CState& s = *static_cast<CState*>(h);
  for (uint64_t b = 0; b < s.nb; ++b) out[b] = s.blocks[b].null ? 1 : 0;
}

uint64_t orc_cs_bytes_history(void* h, uint64_t* out, uint64_t cap) {
  CState& s = *static_cast<CState*>(h);
  for (uint64_t i = 0; i < std::min<uint64_t>(cap, s.bytes_hist.size()); ++i) out[i] = s.bytes_hist[i];
  return s.bytes_hist.size();
}

uint64_t orc_cs_total_bytes(void* h) { return total_bytes(*static_cast<CState*>(h)); }

// sampling_blockwise (src/cppsim/compressed_state.hpp:1600-1681), one
// process: sequential block masses of the decoded blocks, binomial split of
// the shots from mt19937_64(seed), per-block uniforms from
// mt19937_64(seed + 1) against the sequential block CDF (lower_bound), and a
// final shuffle with the first generator.  Returns -2 when the state has no
// finite positive mass, else 0.
int orc_cs_sampling_blockwise(void* h, uint32_t n_shots, uint32_t seed, uint64_t* out) {
  CState& s = *static_cast<CState*>(h);
  std::vector<cplx> blk(s.namps);
  std::vector<double> mass(s.nb, 0.0);
  for (uint64_t b = 0; b < s.nb; ++b) {
    if (s.blocks[b].null) continue;
    decode_block(s.blocks[b], s.namps, blk.data());
    double m = 0.0;
    for (uint64_t j = 0; j < s.namps; ++j) m += std::norm(blk[j]);
    mass[b] = m;
  }
  double total = 0.0;
  for (double m : mass) total += m;
  if (!(total > 0.0) || !std::isfinite(total)) return -2;
  std::mt19937_64 rng(seed);
  std::vector<uint32_t> shots(mass.size(), 0);
  uint32_t left = n_shots;
  double rest = total;
  for (size_t g = 0; g < mass.size() && left > 0; g++) {
    double p = rest > 0 ? std::min(1.0, mass[g] / rest) : 0.0;
    std::binomial_distribution<uint32_t> bin(left, p);
    uint32_t k = (g + 1 == mass.size()) ? left : bin(rng);
    shots[g] = k;
    left -= k;
    rest -= mass[g];
  }
  std::vector<uint64_t> mine;
  std::mt19937_64 rng_local((uint64_t)seed + 1);
  std::uniform_real_distribution<double> uni(0.0, 1.0);
  std::vector<double> cdf(s.namps);
  for (uint64_t b = 0; b < s.nb; b++) {
    uint32_t k = shots[b];
    if (k == 0) continue;
    decode_block(s.blocks[b], s.namps, blk.data());
    double acc = 0.0;
    for (uint64_t j = 0; j < s.namps; j++) {
      acc += std::norm(blk[j]);
      cdf[j] = acc;
    }
    for (uint32_t t = 0; t < k; t++) {
      double r = uni(rng_local) * acc;
      uint64_t j = std::lower_bound(cdf.begin(), cdf.end(), r) - cdf.begin();
      if (j >= s.namps) j = s.namps - 1;
      mine.push_back(b * s.namps + j);
    }
  }
  std::shuffle(mine.begin(), mine.end(), rng);
  std::copy(mine.begin(), mine.end(), out);
  return 0;
}

// Adjoint gradient: psi = U|init>, lambda = 2 H psi, reverse sweep.
void orc_backprop(uint32_t n, const double* init, const void* gates, uint64_t ng, const double* pool, uint64_t nt,
                  const uint64_t* xm, const uint64_t* zm, const uint32_t* ny, const double* cr, const double* ci,
                  uint32_t nparams, double* grads, double* energy) {
  const uint64_t dim = 1ull << n;
  auto gs = load_gates(static_cast<const GateRec*>(gates), ng, pool);
  std::vector<cplx> psi(dim), lam(dim), A(dim);
  std::memcpy(psi.data(), init, dim * 16);
  for (auto& g : gs) apply_dense(psi.data(), n, g.r, g.m.data());
  const Terms T{nt, xm, zm, ny, cr, ci};
  *energy = expectation(psi.data(), dim, T).real();
  apply_pauli_sum(psi.data(), lam.data(), dim, T);
  for (auto& v : lam) v *= 2.0;
  for (uint32_t p = 0; p < nparams; ++p) grads[p] = 0.0;
  for (int64_t i = (int64_t)gs.size() - 1; i >= 0; --i) {
    const Gate& g = gs[i];
    if ((g.r.flags & 1u) && g.r.param_index >= 0) {
      A = psi;
      const Gate R = rotation_of(g, M_PI);
      apply_dense(A.data(), n, R.r, R.m.data());
      double re, im;
      orc_inner(reinterpret_cast<const double*>(lam.data()), reinterpret_cast<const double*>(A.data()), dim, &re, &im);
      grads[g.r.param_index] += re / 2.0;
    }
    const Gate inv = inverse_of(g);
    apply_dense(lam.data(), n, inv.r, inv.m.data());
    apply_dense(psi.data(), n, inv.r, inv.m.data());
  }
}

}  // extern "C"
