/*
 * qsim.h — C ABI of libqsim_hip.so, the MI355X (gfx950) compressed
 * state-vector stage loop.
 *
 * Every entry point returns 0 on success and a negative status on failure;
 * qsim_last_error() then returns the message of the failing call on the
 * calling thread.  Handles are opaque; the library owns every device
 * allocation behind them.  Calls on one handle are not reentrant.  All
 * amplitudes are complex128 stored as interleaved (re, im) doubles; qubit q is
 * bit q of the amplitude index (Qulacs convention).
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   QuantumCircuitOptimizer::optimize        src/cppsim/circuit_optimizer.cpp:113-168
 *   QuantumState / update_quantum_state      src/cppsim/state.hpp:469-482,
 *                                            src/cppsim/circuit.cpp:23-34
 *   CompressedQuantumState                   src/cppsim/compressed_state.hpp:87-1707
 *   pybind surface                           python/cppsim_wrapper.cpp:447-519
 */
#ifndef QSIM_H
#define QSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QSIM_MAX_GATE_QUBITS 8

/* target commutation flags (Qulacs FLAG_X/Y/Z_COMMUTE) */
#define QSIM_COMMUTE_X 1u
#define QSIM_COMMUTE_Y 2u
#define QSIM_COMMUTE_Z 4u

/* gate flags */
#define QSIM_GATE_PARAMETRIC 1u

/*
 * One gate in QuantumGateMatrix form (src/cppsim/gate_matrix.cpp:19-250):
 * a 2^k x 2^k row-major complex matrix whose row/column index bit i belongs to
 * targets[i], applied where every control qubit has its control value.  The
 * matrix lives in a caller-owned pool of interleaved doubles at complex offset
 * matrix_offset.  For parametric rotations (src/vqcsim/parametric_gate.hpp:
 * 102-200) param_index names the parameter and pauli_ids[i] (1=X,2=Y,3=Z) the
 * rotation axis on targets[i]; the matrix then holds exp(+i*theta/2*P).
 */
typedef struct qsim_gate {
  uint32_t n_targets;
  uint32_t n_controls;
  uint32_t targets[QSIM_MAX_GATE_QUBITS];
  uint32_t controls[QSIM_MAX_GATE_QUBITS];
  uint32_t control_values[QSIM_MAX_GATE_QUBITS];
  uint32_t target_commute[QSIM_MAX_GATE_QUBITS];
  uint32_t flags;
  int32_t param_index;
  uint32_t pauli_ids[QSIM_MAX_GATE_QUBITS];
  uint64_t matrix_offset;
} qsim_gate;

const char* qsim_last_error(void);
int qsim_version(void);
int qsim_device_count(int* count);

/* ---- circuit optimizer ------------------------------------------------
 * Greedy pairwise fusion into dense gates of at most max_block_size qubits
 * (replaces QuantumCircuitOptimizer::optimize(circuit, block_size, 0),
 * src/cppsim/circuit_optimizer.cpp:113-168, and gate::merge,
 * src/cppsim/gate_merge.cpp:316-356).  Parametric gates are never merged.
 * out must hold n_gates entries; pool_out must hold
 * qsim_optimize_pool_bound(...) doubles.  Output matrix offsets index pool_out.
 * origin (nullable, n_gates entries) receives the input index of every output
 * gate that passed through unmerged, -1 for merged gates. */
uint64_t qsim_optimize_pool_bound(uint64_t n_gates, uint64_t pool_len, uint32_t max_block_size);
int qsim_optimize(const qsim_gate* gates, uint64_t n_gates, const double* pool, uint64_t pool_len,
                  uint32_t max_block_size, qsim_gate* out, uint64_t* n_out, double* pool_out,
                  uint64_t* pool_out_len, int64_t* origin);

/* ---- uncompressed state in HBM ----------------------------------------
 * QuantumState(n) (src/cppsim/state.hpp:469-482) with the gate sweep of
 * QuantumCircuit::update_quantum_state (src/cppsim/circuit.cpp:23-34) fused
 * into LDS-tiled passes. */
typedef struct qsim_sv qsim_sv;
int qsim_sv_create(uint32_t n_qubits, qsim_sv** out);
int qsim_sv_destroy(qsim_sv* sv);
int qsim_sv_set_basis(qsim_sv* sv, uint64_t index);
int qsim_sv_load(qsim_sv* sv, const double* host_amplitudes);
int qsim_sv_copy(qsim_sv* dst, const qsim_sv* src);
int qsim_sv_read(const qsim_sv* sv, double* host_amplitudes);
int qsim_sv_apply(qsim_sv* sv, const qsim_gate* gates, uint64_t n_gates, const double* pool, uint64_t pool_len);
int qsim_sv_norm_squared(const qsim_sv* sv, double* out);
int qsim_sv_normalize(qsim_sv* sv, double squared_norm);
int qsim_sv_scale(qsim_sv* sv, double factor);
int qsim_sv_inner_product(const qsim_sv* bra, const qsim_sv* ket, double* re, double* im);
/* sum_t coef_t <psi|P_t|psi>; P_t given by X-flip mask, Z-phase mask (Y sets
 * both bits) and its Y count (src/cppsim/pauli_operator.cpp:132-183). */
int qsim_sv_expectation_pauli(const qsim_sv* sv, uint64_t n_terms, const uint64_t* x_masks,
                              const uint64_t* z_masks, const uint32_t* y_counts, const double* coef_re,
                              const double* coef_im, double* out_re, double* out_im);
/* qsim_sv_expectation_pauli without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * term tables, block partials and final sums live in one static device array
 * of the library (104 KiB per device, held until the process ends), so at
 * most 1024 terms are accepted (status -2 beyond).  Results equal
 * qsim_sv_expectation_pauli's bit for bit.  Every call on a device shares
 * that array: calls that overlap on one device race. */
int qsim_sv_expectation_pauli_noalloc(const qsim_sv* sv, uint64_t n_terms, const uint64_t* x_masks,
                                      const uint64_t* z_masks, const uint32_t* y_counts, const double* coef_re,
                                      const double* coef_im, double* out_re, double* out_im);
/* out = sum_t coef_t P_t |psi> (GeneralQuantumOperator::apply_to_state) */
int qsim_sv_apply_pauli_sum(const qsim_sv* sv, qsim_sv* out, uint64_t n_terms, const uint64_t* x_masks,
                            const uint64_t* z_masks, const uint32_t* y_counts, const double* coef_re,
                            const double* coef_im);
/* Adjoint gradient of <H> (ParametricQuantumCircuit::backprop,
 * src/vqcsim/parametric_circuit.cpp:281-308, with backprop_inner_product
 * :180-279).  On entry `sv` holds the initial state (the reference starts from
 * |0...0>); on return it holds U|initial>.  `energy` receives Re<H> and
 * grads[p] the derivative by parameter p of every parametric gate (Pauli
 * rotations exp(+i theta P / 2), as ParametricRX/RY/RZ/PauliRotation; other
 * parametric gates are rejected like the reference's NotImplementedException).
 * H uses the Pauli-term convention of qsim_sv_expectation_pauli.  Runs the
 * forward pass, lambda = 2 H psi in one sweep, then the reverse sweep in fused
 * passes over psi and lambda together. */
int qsim_sv_backprop(qsim_sv* sv, const qsim_gate* gates, uint64_t n_gates, const double* pool,
                     uint64_t pool_len, uint64_t n_terms, const uint64_t* x_masks, const uint64_t* z_masks,
                     const uint32_t* y_counts, const double* coef_re, const double* coef_im,
                     uint64_t n_params, double* energy, double* grads);

void* qsim_sv_device_ptr(qsim_sv* sv);
uint32_t qsim_sv_qubit_count(const qsim_sv* sv);

/* ---- compressed block store in HBM ------------------------------------
 * CompressedQuantumState(n, block_size) (src/cppsim/compressed_state.hpp:
 * 87-239): 2^(n-B) blocks of 2^B amplitudes, each either null (all zero) or
 * held by the error-bounded codec with |x^ - x| <= eps_rel*|x| + eps_abs per
 * double (contract of src/cppsim/compand_bitcomp_pipeline.cuh:14-62). */
#define QSIM_ZERO_EXACT 0     /* GPU semantics, cuda_complex_helper.cuh:96-106 */
#define QSIM_ZERO_THRESHOLD 1 /* CPU semantics, compressed_state.hpp:939-951 */

typedef struct qsim_cs_config {
  double eps_rel;          /* BITCOMP_REL_EB (>= 0; 0 = pure absolute bound) */
  double eps_abs;          /* BITCOMP_ABS_EB */
  int32_t guard_bits;      /* extra quantisation bits below the bound */
  int32_t zero_mode;       /* QSIM_ZERO_EXACT or QSIM_ZERO_THRESHOLD */
  double zero_threshold;   /* ALL_ZERO_THRESHOLD for QSIM_ZERO_THRESHOLD */
  uint64_t scratch_bytes;  /* HBM staging for multi-pass stages (0 = default) */
  uint32_t tile_bits;      /* LDS tile log2 amplitudes (0 = default) */
  uint32_t store_mode;     /* QSIM_STORE_DOUBLE (default) or QSIM_STORE_REGIONS */
  uint64_t hbm_budget;     /* device bytes the store may hold: descriptor tables, staging
                              buffer and arenas (region store: mapped units); 0 = no
                              limit.  Growth past it fails with -1 (RuntimeError); the
                              double-buffered store keeps its state, the region store's
                              state must be reloaded. */
} qsim_cs_config;

/* Payload store of a compressed state (no reference counterpart: the
 * reference keeps one host buffer per block).
 * QSIM_STORE_DOUBLE: two whole arenas, a stage reads one and writes the
 *   other (peak about twice the compressed state).
 * QSIM_STORE_REGIONS: the payload lives in up to 16 regions of a virtual
 *   address reservation, backed by physical units mapped on demand.  A stage
 *   consumes its input region by region and routes its output into the
 *   regions the next stage will consume; each input region's units are
 *   unmapped and reused as soon as it has been consumed (peak about one
 *   compressed state plus one region). */
#define QSIM_STORE_DOUBLE 0
#define QSIM_STORE_REGIONS 1

typedef struct qsim_cs qsim_cs;
int qsim_cs_default_config(qsim_cs_config* cfg);
int qsim_cs_create(uint32_t n_qubits, uint32_t block_size, const qsim_cs_config* cfg, qsim_cs** out);
int qsim_cs_destroy(qsim_cs* cs);
/* partition_and_register_circuit (compressed_state.hpp:673-726) */
int qsim_cs_partition_and_register(qsim_cs* cs, const qsim_gate* gates, uint64_t n_gates, const double* pool,
                                   uint64_t pool_len, uint32_t inner_size);
/* Commutation-aware stage packing (B-aware reordering; no reference
 * counterpart): stages take, in circuit order, every remaining gate whose
 * qubits at or above block_size fit the stage's high set (at most
 * max(inner_size, 2)) and that commutes with every earlier gate left for a
 * later stage.  qsim_partition_commuting only plans (host code, no GPU):
 * stage_of_gate[i] receives gate i's stage.  The _and_register form plans and
 * registers like qsim_cs_partition_and_register. */
int qsim_partition_commuting(uint32_t n_qubits, uint32_t block_size, uint32_t inner_size, const qsim_gate* gates,
                             uint64_t n_gates, const double* pool, uint64_t pool_len, uint32_t* stage_of_gate,
                             uint32_t* n_stages);
int qsim_cs_partition_commuting_and_register(qsim_cs* cs, const qsim_gate* gates, uint64_t n_gates,
                                             const double* pool, uint64_t pool_len, uint32_t inner_size);
/* register_circuit_stages (compressed_state.hpp:464-507): stage s holds gate
 * ids stage_gate_ids[stage_offsets[s] .. stage_offsets[s+1]); block_to_group
 * holds n_stages x 2^(n-B) group ids; num_groups[s] groups per stage. */
int qsim_cs_register_stages(qsim_cs* cs, const qsim_gate* gates, uint64_t n_gates, const double* pool,
                            uint64_t pool_len, uint32_t n_stages, const uint32_t* stage_offsets,
                            const uint32_t* stage_gate_ids, const uint32_t* block_to_group,
                            const uint32_t* num_groups);
int qsim_cs_num_stages(const qsim_cs* cs, uint32_t* out);
int qsim_cs_stage_info(const qsim_cs* cs, uint32_t stage, uint32_t* n_gates, uint32_t* gate_ids,
                       uint32_t* num_groups, uint32_t* block_to_group, uint32_t* n_passes);
/* apply_circuit (compressed_state.hpp:840-937) */
int qsim_cs_apply_circuit(qsim_cs* cs);
int qsim_cs_run_stages(qsim_cs* cs, uint32_t stage_begin, uint32_t stage_end);
/* get_decompressed_state_vector (compressed_state.hpp:1683-1707) */
int qsim_cs_read_vector(qsim_cs* cs, double* host_amplitudes, int normalise);
int qsim_cs_read_block(qsim_cs* cs, uint64_t block, double* host_amplitudes);
int qsim_cs_set_basis(qsim_cs* cs, uint64_t index);
int qsim_cs_load(qsim_cs* cs, const double* host_amplitudes);
/* Compressed-store checkpoint files (no reference counterpart; the
 * checkpoint format of a long run, so that it can resume), for both payload
 * stores: a header with the store kind, shape and codec settings, the payload
 * layout (double store: sub-arena table and allocation cursors; region
 * store: the region key, span and per-region extents), the null flags, the
 * chunk descriptors and the payload words.  next_stage is stored for the
 * caller (the first stage still to run).  Loading requires a store of the
 * same kind, qubit count, block size and codec settings (status -2
 * otherwise), validates every descriptor against the payload extents (-2 on
 * a corrupt file) and the file size against the layout (-1 when truncated)
 * before it changes the store; an I/O failure while the payload is copied
 * resets the store to |0...0> (-1). */
int qsim_cs_save_checkpoint(qsim_cs* cs, const char* path, uint64_t next_stage);
int qsim_cs_load_checkpoint(qsim_cs* cs, const char* path, uint64_t* next_stage);
int qsim_cs_null_bitmap(const qsim_cs* cs, uint8_t* out);
/* sampling_blockwise (compressed_state.hpp:1600-1681): n_shots basis indices
 * into out, drawn with seed as the reference does in one process (block
 * masses, binomial split over blocks from mt19937_64(seed), per-block
 * uniforms from mt19937_64(seed + 1), shuffle).  A state without finite
 * positive mass fails with status -2 (the reference's std::domain_error). */
int qsim_cs_sampling_blockwise(qsim_cs* cs, uint32_t n_shots, uint32_t seed, uint64_t* out);
int qsim_cs_block_bytes(qsim_cs* cs, uint64_t* out);
/* Total compressed bytes after the initial state and after every stage
 * (append_total_SV_bytes, compressed_state.hpp:374-437); history[0] is the
 * initial state.  *n receives the history length. */
int qsim_cs_bytes_history(const qsim_cs* cs, uint64_t* out, uint64_t cap, uint64_t* n);
int qsim_cs_max_sv_bytes(const qsim_cs* cs, uint64_t* out);
/* Device memory of the payload store (no reference counterpart): out3[0] the
 * arena bytes held now (region store: mapped units), out3[1] the most held
 * at once, out3[2] the descriptor tables (double store: input and output;
 * region store: one table, encoded in place, plus the copy of a stage
 * batch's input descriptors). */
int qsim_cs_store_bytes(const qsim_cs* cs, uint64_t* out3);
/* Per-stage measurement of the last run: for each executed stage
 * {gpu_seconds, bytes_in, bytes_out, decompressed_amplitudes, n_passes}. */
int qsim_cs_stage_stats(const qsim_cs* cs, double* out5, uint64_t cap, uint64_t* n);
#define QSIM_TIMER_COUNT 10 /* decomp gate comp fused_swap allgather hmalloc dmalloc h2d d2h dfree */
/* get_total_times (python/cppsim_wrapper.cpp:503-519), seconds since the
 * store was created: decomp, gate and comp from HIP events around the stage
 * launches (DEC->RAW passes, RAW->RAW passes plus the fused DEC->ENC launch
 * of a single-pass stage, RAW->ENC passes); dmalloc and dfree the wall time
 * of device allocation, release, region mapping and unmapping; h2d and d2h
 * the wall time of amplitude copies in load and readout.  fused_swap,
 * allgather and hmalloc are 0 (one GPU, no host block buffers). */
int qsim_cs_times(const qsim_cs* cs, double* out);
/* sum_t coef_t <psi|P_t|psi> over the compressed store (no reference
 * counterpart: the reference reads expectation values from the decompressed
 * vector).  Blocks are decoded in batches; X/Y factors must lie on qubits
 * inside one decode batch (the block qubits and the low block-id bits). */
int qsim_cs_expectation_pauli(qsim_cs* cs, uint64_t n_terms, const uint64_t* x_masks, const uint64_t* z_masks,
                              const uint32_t* y_counts, const double* coef_re, const double* coef_im,
                              double* out_re, double* out_im);
int qsim_cs_norm_squared(qsim_cs* cs, double* out);
int qsim_cs_inner_product_sv(qsim_cs* cs, const qsim_sv* sv, double* re, double* im);
uint32_t qsim_cs_qubit_count(const qsim_cs* cs);
uint32_t qsim_cs_block_size(const qsim_cs* cs);

/* Adjoint with compressed forward checkpoints (the forward_with_checkpoints(k)
 * / adjoint(params, obs, grad_out) pair; the reference has no counterpart, it
 * extends ParametricQuantumCircuit::backprop, parametric_circuit.cpp:281-308).
 * qsim_adj_forward applies the circuit to `sv` and stores, in compressed
 * stores of `block_size` with the codec settings `codec`, the state at the
 * start of every segment of `segment_gates` gates.  qsim_adj_backward then
 * forms lambda = 2 H psi from `sv` (which must still hold the forward
 * result), and for each segment from the last restores its checkpoint,
 * recomputes the segment and sweeps it backwards with lambda: energy and
 * gradients as qsim_sv_backprop, with drift confined to one segment.
 * A handle keeps its checkpoint stores between qsim_adj_forward calls with
 * the same qubit count, block size and codec settings, so a repeated forward
 * allocates nothing. */
typedef struct qsim_adj qsim_adj;
int qsim_adj_create(qsim_adj** out);
int qsim_adj_destroy(qsim_adj* adj);
int qsim_adj_forward(qsim_adj* adj, qsim_sv* sv, const qsim_gate* gates, uint64_t n_gates, const double* pool,
                     uint64_t pool_len, uint64_t segment_gates, uint32_t block_size, const qsim_cs_config* codec);
int qsim_adj_backward(qsim_adj* adj, qsim_sv* sv, uint64_t n_terms, const uint64_t* x_masks, const uint64_t* z_masks,
                      const uint32_t* y_counts, const double* coef_re, const double* coef_im, uint64_t n_params,
                      double* energy, double* grads);
int qsim_adj_checkpoint_bytes(const qsim_adj* adj, uint64_t* n_checkpoints, uint64_t* bytes);

/* ---- synchronisation / timing helpers ---------------------------------- */
int qsim_synchronize(void);

/* ---- batched circuits --------------------------------------------------
 * Many small circuits of uneven qubit counts, each with its own Hermitian
 * Pauli-sum observable: energy and adjoint gradient of every instance (the
 * semantics of qsim_sv_backprop from |0...0>).  qsim_batch_run plans every
 * instance once and, with use_graph, captures the whole batch into one
 * hipGraph on the first call and replays it afterwards.  Instances of 10 to
 * 16 qubits run in one persistent kernel launch (a work-group per instance,
 * each taking the next from a counter; QOMPSIM_BATCH_PERSISTENT=0 gives
 * them launches of their own), larger ones in launches of their own on
 * forked streams.  qsim_batch_tune
 * times the forward engine / tile choices and the adjoint tile sizes per qubit
 * count and keeps the fastest.  No counterpart in the reference, which runs
 * one circuit per call (the batch entry points the boundary adds for this
 * path's batched workload). */
typedef struct qsim_batch qsim_batch;
int qsim_batch_create(qsim_batch** out);
int qsim_batch_destroy(qsim_batch* b);
int qsim_batch_add(qsim_batch* b, uint32_t n_qubits, const qsim_gate* gates, uint64_t n_gates, const double* pool,
                   uint64_t pool_len, uint64_t n_terms, const uint64_t* x_masks, const uint64_t* z_masks,
                   const uint32_t* y_counts, const double* coef_re, const double* coef_im, uint64_t n_params,
                   uint64_t* index);
int qsim_batch_tune(qsim_batch* b, int reps);
/* engines: -1 default, 0 LDS tile (2^tile_bits amplitudes), 1 wave tile */
int qsim_batch_config(const qsim_batch* b, uint32_t n_qubits, int32_t* fwd_engine, uint32_t* fwd_tile_bits,
                      int32_t* adj_engine, uint32_t* adj_tile_bits);
/* energies[i]; grads[i * qsim_batch_max_params(b) + p] */
int qsim_batch_run(qsim_batch* b, int use_graph, double* energies, double* grads);
uint64_t qsim_batch_size(const qsim_batch* b);
uint64_t qsim_batch_max_params(const qsim_batch* b);
/* Algorithmic HBM bytes of one qsim_batch_run (plans the instances first if
 * needed): per instance of n qubits 16 * 2^n for |0...0>, 32 * 2^n per
 * forward pass and for lambda = H psi, 64 * 2^n per reverse pass (psi and
 * lambda).  *persistent receives the number of instances the persistent
 * kernel runs.  No counterpart in the reference: the measurement of the
 * batched path's roofline. */
int qsim_batch_traffic(qsim_batch* b, uint64_t* bytes, uint64_t* persistent);

#ifdef __cplusplus
}
#endif
#endif /* QSIM_H */
