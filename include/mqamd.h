/*
 * mqamd.h - C ABI of the MI355X (gfx950) state-vector backend.
 *
 * The library owns every device buffer; the caller owns every host buffer and no pointer
 * escapes a call.  All entry points return an int status (MQA_OK == 0) and leave a
 * thread-local message readable through mqa_last_error().  Calls on one simulator handle
 * are serialised by the caller (the reference documents the same restriction:
 * ccsrc/include/simulator/vector/vector_state.tpp:715 mutates the rng inside const methods).
 *
 * State layout: interleaved complex (re, im), little-endian qubit order - amplitude index
 * bit q <=> qubit q (reference: ccsrc/lib/simulator/vector/detail/cpu_common/
 * cpu_vector_core_policy.cpp:72).  c64 states hold float pairs, c128 states double pairs.
 *
 * Each entry point names the reference interface it replaces (file:line, relative to the
 * mindquantum 0.12.0 source tree).
 */
#ifndef MQAMD_H_
#define MQAMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (mapped to Python exceptions by the host layer) ---- */
#define MQA_OK 0
#define MQA_ERR_ARG 1         /* invalid argument         -> ValueError   */
#define MQA_ERR_OOM 2         /* device allocation failed -> MemoryError  */
#define MQA_ERR_HIP 3         /* HIP runtime failure      -> RuntimeError */
#define MQA_ERR_UNSUPPORTED 4 /* gate kind not supported  -> ValueError   */
#define MQA_ERR_RUNTIME 5     /* other runtime failure    -> RuntimeError */

/* ---- precision ---- */
#define MQA_C64 0  /* complex64: float amplitudes, float arithmetic  */
#define MQA_C128 1 /* complex128: double amplitudes, double arithmetic */

/* ---- gate kinds (reference: ccsrc/include/ops/gate_id.h:25-70) ---- */
#define MQA_GATE_I 0
#define MQA_GATE_X 1
#define MQA_GATE_Y 2
#define MQA_GATE_Z 3
#define MQA_GATE_H 4
#define MQA_GATE_S 5
#define MQA_GATE_SDAG 6
#define MQA_GATE_T 7
#define MQA_GATE_TDAG 8
#define MQA_GATE_SX 9
#define MQA_GATE_SXDAG 10
#define MQA_GATE_SWAP 11
#define MQA_GATE_ISWAP 12
#define MQA_GATE_ISWAPDAG 13
#define MQA_GATE_MATRIX 14 /* fixed dense 2^k x 2^k matrix, k = n_obj <= MQA_MAX_OBJ */
/* parameterised gates: angle = const_val + sum_t coeff[t] * params[param[t]] */
#define MQA_GATE_RX 32
#define MQA_GATE_RY 33
#define MQA_GATE_RZ 34
#define MQA_GATE_PS 35 /* PhaseShift diag(1, e^{i a}) */
#define MQA_GATE_GP 36 /* GlobalPhase e^{-i a} I      */
#define MQA_GATE_RXX 37
#define MQA_GATE_RYY 38
#define MQA_GATE_RZZ 39
/* parameterised dense matrix (custom gates of one or two parameters, parameterised Rn; 1 or 2 targets):
 * the matrix depends on the parameters through a host function, so its values come per call and per
 * sample from mqa_circuit_set_custom_matrices.  mat_begin is the offset (in doubles) of the gate's
 * block in a sample's row: [G][G^dagger][M_0], each D x D interleaved complex (D = 2^n_obj), with
 * M_0 = dG/dv_0 G^dagger for the gate's first parameter expression v_0 (the angle fields).  A gate
 * with further expressions v_j is followed by one MQA_GATE_PMATRIX_TAP record per expression (same
 * targets and controls; mat_begin: the offset of M_j = dG/dv_j G^dagger; angle fields: v_j), which
 * only contributes a derivative. */
#define MQA_GATE_PMATRIX 40
#define MQA_GATE_PMATRIX_TAP 41

#define MQA_MAX_OBJ 4

/* One gate of a circuit.  64 bytes, no padding surprises across compilers. */
typedef struct mqa_gate {
    int32_t kind;       /* MQA_GATE_* */
    int32_t n_obj;      /* number of target qubits (1..MQA_MAX_OBJ) */
    int32_t obj[6];     /* target qubits; obj[i] <=> bit i of the gate's matrix index */
    uint64_t ctrl_mask; /* control qubits as a bit mask */
    double const_val;   /* constant part of the angle (parameterised kinds) */
    int32_t term_begin; /* first (param, coeff) term of the angle */
    int32_t n_term;     /* number of terms */
    int32_t mat_begin;  /* MATRIX: first complex entry in the matrix pool (row-major) */
    int32_t requires_grad; /* 0: the gate gets no derivative tap */
} mqa_gate;

/* One Pauli term c * X^x Y^y Z^z (masks are disjoint) with a real coefficient
 * (reference: ccsrc/include/core/mq_base_types.h:67-76, mindquantum/core/operators/hamiltonian.py:94-101). */
typedef struct mqa_pauli_term {
    uint64_t x_mask;
    uint64_t y_mask;
    uint64_t z_mask;
    double coeff;
} mqa_pauli_term;

typedef struct mqa_sim mqa_sim;
typedef struct mqa_circuit mqa_circuit;
typedef struct mqa_ham mqa_ham;

/* ---- library ---- */
/* Message of the last failed call on this thread (empty when none). */
const char* mqa_last_error(void);
/* Library version string. */
const char* mqa_version(void);
/* Number of visible HIP devices (0 without a GPU).  Never fails. */
int mqa_device_count(void);
/* Tuning knobs (value < 0 restores the default): "tile_qubits_c64", "tile_qubits_c128",
 * "adj_tile_qubits_c64", "adj_tile_qubits_c128", "group_width", "lowq_c64", "lowq_c128",
 * "max_ops_per_pass" (0, the default: the scheduler keeps the cheapest of a few caps under its
 * pass cost model; 1 = one gate per pass), "hbm_budget_bytes" (also the environment variable
 * MQAMD_HBM_BUDGET_BYTES; 0 = 85% of free memory), "profile", "ham_tile_qubits_c64",
 * "ham_tile_qubits_c128", "ham_tile_engine", "jit" (0 off, 1 background compile, 2 compile before
 * the first launch), "jit_min_amps" (launches of at least this many amplitudes use compiled passes),
 * "jit_min_runs" (8: so do smaller ones once their plan has run that many times; 0: never),
 * "checkpoints", "phase_free" (1, the default: mqa_sim_grad_batch
 * runs plans in which uncontrolled RX / RY / RZ / GP drop their global phase, see there; 0: exact
 * plans), "jit_single_buffer" (1, the default: compiled back-sweep passes keep psi and lambda in
 * registers between register groups and share one LDS tile buffer, so twice as many blocks fit a CU;
 * 0: two tile buffers), "dense_mfma" (1, the default: complex64 5-qubit blocks of
 * mqa_sim_apply_dense_blocks run on the FP32 matrix cores; 0: on the vector ALUs), "phase_ops" (1, the
 * default: with "jit" on, the plans of batched gradients and gate application merge RZ, PS,
 * uncontrolled RZZ and the constant phases Z, S, T (their inverses and controlled forms such as CZ)
 * into diagonal phase blocks that need no local qubit; 2: also with "jit" off; 0: one op per gate).
 * "items_jit_qubits" (23: items of an uneven batch with at least this many qubits run on their own
 * compiled passes; 0: every item in the shared interpreted launches).  mqa_get_option also reads the counters "jit_launches", "jit_compile_ms",
 * "grad_recomputes", "checkpoints_used", "phase_free_used" and "jit_single_buffer_launches".  No counterpart
 * in the reference (its CPU kernels have no tuning surface). */
int mqa_set_option(const char* key, int64_t value);
int mqa_get_option(const char* key, int64_t* value);

/* Kernel timing: with the option "profile" set to 1, HIP events are recorded on the
 * simulator's own stream around every launch of the named kernel family
 * ("tile_pass_fwd", "tile_pass_adj", "hamiltonian") and summed here, together with the
 * algorithmic HBM bytes of those launches (each state read and written once per pass).
 * No counterpart in the reference (it has no device timing). */
int mqa_profile_get(const char* kernel, int64_t* launches, double* total_ms, double* bytes);
int mqa_profile_reset(void);
/* Streaming-copy probe: times `reps` launches of a plain element-wise copy kernel (one read and
 * one write of `bytes` each) on fresh device buffers and returns the average launch time;
 * in_place = 1 writes every element back where it was read (the access pattern of a gate pass), 2
 * does so with non-temporal accesses (as the tile passes stream the state).  The achievable HBM
 * rate of this GPU, and the kernel of known traffic that calibrates the HBM byte counters of a
 * profiler.  No counterpart in the reference. */
int mqa_copy_probe(uint64_t bytes, int32_t reps, int32_t in_place, double* ms_per_copy);

/* ---- circuit (reference: mindquantum/core/circuit/circuit.py:1108-1130 get_cpp_obj) ----
 * Host-side object: validated gate list plus the cached tile-pass schedules.
 * term_param/term_coeff hold the linear angle terms referenced by gates; mats holds
 * interleaved complex matrices referenced by MATRIX gates.  n_params bounds param indices.
 * param_requires_grad (n_params flags, NULL = all) drops no-grad parameters from the
 * jacobian, as ParameterResolver::GetRequiresGradParameters does (basic_gate.h:208-242). */
int mqa_circuit_create(const mqa_gate* gates, int32_t n_gates, const int32_t* term_param, const double* term_coeff,
                       int32_t n_terms, const double* mats, int64_t n_mat_entries, int32_t n_params,
                       const uint8_t* param_requires_grad, mqa_circuit** out);
int mqa_circuit_destroy(mqa_circuit* circ);
/* Values of a circuit's parameterised matrices (MQA_GATE_PMATRIX / _TAP) for the calls that follow:
 * rows x stride doubles, row b for sample b of a batched call (rows == 1: one row shared by every
 * sample).  A call that runs the circuit on more samples than rows (other than 1) fails.  The reference
 * evaluates the gate's matrix function per sample inside its C++ loop (CustomGate, vector_state.tpp:
 * 402-413); here the host evaluates it and the kernels read the table. */
int mqa_circuit_set_custom_matrices(mqa_circuit* circ, int64_t rows, int64_t stride, const double* table);
/* Tile width (qubits per LDS tile) of this circuit's forward (adjoint == 0) or back-sweep passes
 * in one precision; tile_qubits < 0 returns to the global option ("tile_qubits_c64" ...).  Set by
 * the autotuner (mindquantum_amd.autotune), which times candidate widths on the device.  The get
 * call returns the width in effect.  No counterpart in the reference (its CPU kernels have no
 * tiles). */
int mqa_circuit_set_tile_qubits(mqa_circuit* circ, int32_t dtype, int32_t adjoint, int32_t tile_qubits);
int mqa_circuit_get_tile_qubits(const mqa_circuit* circ, int32_t dtype, int32_t adjoint, int32_t* tile_qubits);
/* Export the tile-pass schedule (host only, no GPU needed) for inspection and tests.
 * Layout of out (int64 words): [n_passes, then per pass: local_mask, n_gates, gate ids...].
 * adjoint bit 0 exports the reverse (back-sweep) schedule, bit 1 its phase-free form, bit 4
 * (+ 16) the form with phase blocks (option "phase_ops": RZ, PS, RZZ and constant phase gates
 * merged into diagonal blocks that need no local qubit).  *len gets the words needed. */
int mqa_circuit_schedule(const mqa_circuit* circ, int32_t n_qubits, int32_t dtype, int32_t adjoint, int64_t* out,
                         int64_t cap, int64_t* len);
/* Generate and compile (hipRTC, host only, no GPU needed) the per-circuit kernels of the
 * forward (adjoint == 0) or back-sweep (1) schedule, or of their phase-free forms (2, 3: those
 * the hermitian gradient runs); + 8 for the single-buffer form of a back-sweep (option
 * "jit_single_buffer"), + 16 for the schedule with phase blocks (option "phase_ops"); as the library
 * does in the background for large launches (option "jit").  *seconds gets the compile wall time.  No counterpart in the
 * reference: a build check of the code generator for tests and for callers that want to pay
 * the compile cost up front. */
int mqa_circuit_jit_compile(const mqa_circuit* circ, int32_t n_qubits, int32_t dtype, int32_t adjoint,
                            double* seconds);

/* ---- Hamiltonian (reference: ccsrc/include/ops/hamiltonian.h:30-55, how_to_ == ORIGIN) ---- */
int mqa_ham_create(const mqa_pauli_term* terms, int32_t n_terms, mqa_ham** out);
/* Sparse Hamiltonian from a dim x dim CSR matrix (hamiltonian.py:59-76, how_to_ == FRONTEND;
 * Csr_Dot_Vec / ExpectationOfCsr, cpu_vector_core_dot_like.cpp:132-225): indptr [dim + 1],
 * indices [nnz] (< dim), data [nnz] interleaved complex, in the little-endian amplitude order.  dim
 * must equal 2^n of the simulators it is used with.  The operator need not be hermitian: the
 * non-hermitian gradient applies its conjugate transpose where the reference uses herm_hams.
 * Not accepted by mqa_grad_items. */
int mqa_ham_create_csr(uint64_t dim, uint64_t nnz, const int64_t* indptr, const int64_t* indices, const double* data,
                       mqa_ham** out);
int mqa_ham_destroy(mqa_ham* ham);

/* ---- simulator (reference: ccsrc/python/simulator/include/python/vector/bind_vec_state.h:46-84, BindSim) ---- */
/* VectorState(n_qubits, seed) (vector_state.tpp:63-68); the state starts as |0...0>. */
int mqa_sim_create(int32_t n_qubits, int32_t dtype, uint32_t seed, mqa_sim** out);
int mqa_sim_destroy(mqa_sim* sim);
/* reset() (vector_state.tpp Reset): back to |0...0>. */
int mqa_sim_reset(mqa_sim* sim);
/* copy() (bind_vec_state.h): deep copy of the device state. */
int mqa_sim_copy(const mqa_sim* sim, mqa_sim** out);
int mqa_sim_n_qubits(const mqa_sim* sim, int32_t* n_qubits, int32_t* dtype);
/* get_qs() (cpu_vector_core_policy.cpp:99-118): device -> host, dim = 2^n interleaved pairs. */
int mqa_sim_get_qs(const mqa_sim* sim, void* host_out, uint64_t dim);
/* set_qs() (cpu_vector_core_policy.cpp:120-130): host -> device; the length must equal 2^n
 * ("state size not match").  No normalisation here: mqsim.py:537-544 normalises in Python. */
int mqa_sim_set_qs(mqa_sim* sim, const void* host_in, uint64_t dim);
/* apply_circuit(circ, pr) (vector_state.tpp:948-960); params holds the circuit's parameters. */
int mqa_sim_apply_circuit(mqa_sim* sim, const mqa_circuit* circ, const double* params, int32_t n_params);
/* A sequence of dense k-qubit matrices (k = 1..5), each applied in one streaming sweep over the
 * state: the apply of ApplyNQubitsMatrix (cpu_vector_core_matrix_gate.cpp:28-83, no controls) for the
 * dense blocks a gate-fusion pass builds from runs of gates (mindquantum_amd.fusion).  Block b acts
 * on block_qubits[b] qubits, listed in turn in `qubits` (matrix index bit i <-> the block's i-th
 * qubit), with a row-major 2^k x 2^k complex matrix of interleaved doubles, listed in turn in
 * `matrices`.  Every block is checked before the first sweep runs. */
int mqa_sim_apply_dense_blocks(mqa_sim* sim, int32_t n_blocks, const int32_t* block_qubits, const int32_t* qubits,
                               const double* matrices);
/* apply_hamiltonian(ham) (vector_state.tpp:962-974): state <- H state. */
int mqa_sim_apply_hamiltonian(mqa_sim* sim, const mqa_ham* ham);
/* get_expectation(ham, circ, pr) (vector_state.tpp:711-726): <0|U^dag H U|0> on a copy of the
 * state; the simulator's own state is left unchanged.  circ may be NULL (no circuit). */
int mqa_sim_expectation(const mqa_sim* sim, const mqa_ham* ham, const mqa_circuit* circ, const double* params,
                        int32_t n_params, double out[2]);
/* mqa_sim_expectation with no circuit, <state|H|state> of the simulator's own state, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  It calls no hipMalloc / hipFree: the
 * Hamiltonian's tables, the per-block partials and the reduced sum live in one file-scope device
 * array of 4 MiB per GPU (held until the process ends); Pauli-term Hamiltonians of at most 8192 flip
 * groups and 60000 terms (MQA_ERR_ARG beyond; CSR matrices are rejected).  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (from several threads or streams) race.
 * Same kernels and result as mqa_sim_expectation(sim, ham, NULL, ...). */
int mqa_sim_expectation_noalloc(const mqa_sim* sim, const mqa_ham* ham, double out[2]);
/* get_expectation(ham, circ_right, circ_left, sim_left, pr) (vector_state.tpp:728-768):
 * <sim_left|Ul^dag H Ur|sim>.  left may be NULL (== sim); either circuit may be NULL. */
int mqa_sim_expectation_lr(const mqa_sim* sim, const mqa_sim* left, const mqa_ham* ham, const mqa_circuit* circ_right,
                           const mqa_circuit* circ_left, const double* params, int32_t n_params, double out[2]);
/* get_expectation_with_grad_multi_multi (vector_state.tpp:1020-1028 -> gradient.h:100-121):
 * hermitian adjoint gradient for a batch of encoder rows sharing one ansatz vector.
 * Param index p < n_enc reads enc[b*n_enc + p]; p >= n_enc reads ans[p - n_enc].
 * f_out: [batch][n_hams][2] (complex <psi|H|psi>); g_out: [batch][n_hams][n_enc+n_ans]
 * (real; the reference returns 2*Re with a zero imaginary part, gradient.h:124-133).
 * Samples run in chunks that fit the HBM budget (two states per sample, three when several
 * hamiltonians share a kept copy of the forward state; if that copy does not fit, the forward
 * sweep is recomputed per hamiltonian).  With the option "checkpoints" = c > 0 the forward sweep
 * also stores psi at up to c back-sweep pass boundaries (as many as the budget holds for one
 * sample, one state each; read back with the counter "checkpoints_used") and the back-sweep
 * continues from each stored state instead of the uncomputed one, bounding the rounding drift of
 * the uncompute; the forward sweep is then recomputed per hamiltonian.  With the option
 * "phase_free" (default 1) the forward sweep, the back-sweep and the checkpoints drop the global
 * phases of uncontrolled RX / RY / RZ / GP gates alike (neither <psi|H|psi> nor the taps
 * <lambda|dG|psi> see a phase common to psi and lambda): RZ becomes a phase on one amplitude,
 * RX / RY rotate as three shears, and the RZ and GP taps use the fact that <lambda|psi> =
 * <psi|H|psi> is real, so they run only when every observable is hermitian (Pauli terms always
 * are; a CSR matrix must equal its conjugate transpose); the counter "phase_free_used" tells
 * whether the last call ran them. */
int mqa_sim_grad_batch(const mqa_sim* sim, const mqa_ham* const* hams, int32_t n_hams, const mqa_circuit* circ,
                       const double* enc, int64_t batch, int32_t n_enc, const double* ans, int32_t n_ans, double* f_out,
                       double* g_out);
/* How mqa_sim_grad_batch sizes its device workspaces (host only, no GPU): for a workspace budget
 * (85% of free memory plus the workspaces the simulator already holds), one state of state_bytes,
 * side_bytes of small per-sample tables, n_hams observables and n_checkpoints checkpoints, *chunk
 * samples run per chunk, *keep tells whether the forward state is kept for further observables, and
 * need[] holds the bytes of the four workspaces (psi, lambda, kept state, checkpoints).  held[] are
 * their current sizes; release[i] = 1 when workspace i is freed before any allocation (it must
 * grow, is no longer used, or a larger-than-needed buffer would not fit beside the others).  No
 * counterpart in the reference (its workspaces are host memory per thread). */
int mqa_grad_workspace_plan(uint64_t budget_bytes, uint64_t state_bytes, uint64_t side_bytes, int32_t n_hams,
                            int32_t n_checkpoints, int64_t batch, const uint64_t held[4], int64_t* chunk, int32_t* keep,
                            uint64_t need[4], int32_t release[4]);
/* get_expectation_with_grad_non_hermitian_multi_multi (vector_state.tpp:998-1008 ->
 * gradient.h:83-98, 199-268): f = <phi_l|H|psi_r> with psi_r = circ_right |sim>, phi_l =
 * circ_left |left> (left NULL: sim), and its complex derivative df/dp = <phi_l|H|d psi_r> +
 * <d phi_l|H|psi_r> for every parameter (both circuits share the parameter list; no hermitian
 * doubling).  f_out: [batch][n_hams][2]; g_out: [batch][n_hams][n_enc+n_ans][2] (re, im).  The
 * hamiltonians have real coefficients (hermitian), as everywhere in this ABI. */
int mqa_sim_grad_nonherm(const mqa_sim* sim, const mqa_sim* left, const mqa_ham* const* hams, int32_t n_hams,
                         const mqa_circuit* circ_right, const mqa_circuit* circ_left, const double* enc, int64_t batch,
                         int32_t n_enc, const double* ans, int32_t n_ans, double* f_out, double* g_out);
/* Expectation and hermitian adjoint gradient of many independent circuits of uneven size in
 * shared launches (one launch per sweep step over the tiles of every circuit).  Item i: circuit
 * circs[i] on n_qubits[i] qubits starting from |0...0>, hamiltonian hams[i], parameter values
 * params[param_offset[i] ..] (circs[i]'s own parameter order).  Outputs: f_out[2i], f_out[2i+1]
 * (<psi|H|psi>), g_out[param_offset[i] ..] (real gradient).  Items run in chunks that fit the HBM
 * budget.  The reference has no batched entry for different circuits: a caller makes one
 * get_expectation_with_grad call per circuit and simulator (mqsim.py:255-439); this is the same
 * result for a whole list.  With "jit" on, items of at least "items_jit_qubits" qubits (23 by
 * default; 0: none) run one by one as the batched gradient of one sample on their per-circuit
 * compiled passes (a launch of one such state fills the device), the rest in the shared launches. */
int mqa_grad_items(int32_t n_items, const mqa_circuit* const* circs, const mqa_ham* const* hams,
                   const int32_t* n_qubits, const double* params, const int64_t* param_offset, int32_t dtype,
                   double* f_out, double* g_out);
/* sampling_measure_ending_without_noise (vector_state.tpp:1326-1373): applies circ (the circuit
 * without its measurements; NULL: none) to a copy of the state and returns, for each of the `shots`
 * sorted uniform samples u in [0, 1), the amplitude index i with C(i-1) < u * C(dim-1) <= C(i),
 * C the cumulative probability (lower bound).  The caller draws the uniforms (the reference draws
 * them from its seeded mt19937; results agree in distribution, not sample by sample) and reads the
 * measured qubits' bits from the indices.  out_index: [shots]. */
int mqa_sim_sample(const mqa_sim* sim, const mqa_circuit* circ, const double* params, int32_t n_params,
                   const double* uniforms, uint64_t shots, uint64_t* out_index);
/* Expectations of a batch of angle settings of one circuit (the evaluations of
 * get_expectation_with_grad_parameter_shift_multi_multi, vector_state.tpp:1033-1300): sample b
 * runs circ from the simulator's state with gate angle slot j set to angles[b * n_slots + j] (slots
 * in the circuit's order of parameterised gates, the angle as the gate's matrix takes it), and
 * f_out[b][h] = <psi_b|H_h|psi_b> (re, im).  Samples run in chunks that fit the HBM budget. */
int mqa_sim_expectation_angles(const mqa_sim* sim, const mqa_ham* const* hams, int32_t n_hams,
                               const mqa_circuit* circ, const double* angles, int64_t batch, double* f_out);
/* complex64() / complex128() (vector_state.tpp:770-775 astype): a new simulator holding the
 * state cast to dtype. */
int mqa_sim_astype(const mqa_sim* sim, int32_t dtype, mqa_sim** out);
/* blas.inner_product(a, b) (bind_vec_state.h:86-90, blas.h): <a|b>. */
int mqa_inner_product(const mqa_sim* bra, const mqa_sim* ket, double out[2]);

#ifdef __cplusplus
}
#endif
#endif /* MQAMD_H_ */
