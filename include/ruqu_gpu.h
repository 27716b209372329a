/*
 * ruqu_gpu.h — C ABI of the MI355X state-vector engine (libruqu_gpu.so).
 *
 * Drop-in boundary for ruqu-core's circuit/state API.  Each entry point names
 * the Rust item it replaces (paths relative to the ruvnet/ruqu checkout).
 * Plain pointers and sizes only; every function returns an rq_status and never
 * aborts.  rq_last_error() gives the message of the last failure on the
 * calling thread.
 *
 * Conventions (identical to the reference):
 *   - qubit q is bit q of the amplitude index (little-endian; state.rs:128-136)
 *   - amplitudes are interleaved (re, im) pairs (types.rs:8-11)
 *   - 2q matrix row/column index = q1_bit*2 + q2_bit (gate.rs:150-153)
 *   - uniform draws in [0,1) are inputs: the reference draws them from its
 *     own StdRng (state.rs:302); measurement here consumes caller draws.
 *
 * Threading: a state handle owns one HIP stream and must be used by one
 * thread at a time; distinct handles may be used from distinct threads.
 */
#ifndef RUQU_GPU_H
#define RUQU_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* types.rs:8-11 (f64) and mixed_precision.rs:33-39 (f32). */
typedef struct { double re, im; } rq_c64;
typedef struct { float re, im; } rq_c32;

/* error.rs:8-23 variants in declaration order, then the engine's own. */
typedef enum {
  RQ_OK = 0,
  RQ_ERR_QUBIT_LIMIT = 1,     /* QuantumError::QubitLimitExceeded */
  RQ_ERR_INVALID_QUBIT = 2,   /* QuantumError::InvalidQubitIndex */
  RQ_ERR_ALLOC = 3,           /* QuantumError::MemoryAllocationFailed (hipMalloc failed) */
  RQ_ERR_INVALID_STATE = 4,   /* QuantumError::InvalidStateVector */
  RQ_ERR_CIRCUIT = 5,         /* QuantumError::CircuitError */
  RQ_ERR_DEVICE = 6,          /* HIP runtime failure or no usable gfx950 device */
  RQ_ERR_ARGUMENT = 7         /* null pointer / bad size passed across the ABI */
} rq_status;

/* Precision of a state: complex-f64 (QuantumState) or complex-f32
 * (QuantumStateF32, mixed_precision.rs:193-200). */
typedef enum { RQ_C64 = 0, RQ_C32 = 1 } rq_precision;

/* gate.rs:8-36, declaration order. */
typedef enum {
  RQ_GATE_H = 0, RQ_GATE_X, RQ_GATE_Y, RQ_GATE_Z, RQ_GATE_S, RQ_GATE_SDG,
  RQ_GATE_T, RQ_GATE_TDG, RQ_GATE_RX, RQ_GATE_RY, RQ_GATE_RZ, RQ_GATE_PHASE,
  RQ_GATE_CNOT, RQ_GATE_CZ, RQ_GATE_SWAP, RQ_GATE_RZZ,
  RQ_GATE_MEASURE, RQ_GATE_RESET, RQ_GATE_BARRIER, RQ_GATE_UNITARY1Q
} rq_gate_kind;

/* One `Gate` value (gate.rs:8-36): q0 is the (first) qubit, q1 the second
 * qubit of a 2q gate (CNOT: control=q0, target=q1), theta the angle of
 * Rx/Ry/Rz/Phase/Rzz, u the matrix of Unitary1Q. 88 bytes. */
typedef struct {
  int32_t kind;
  uint32_t q0, q1, _pad;
  double theta;
  rq_c64 u[2][2];
} rq_gate;

/* types.rs:230-234 MeasurementOutcome. */
typedef struct {
  uint32_t qubit;
  int32_t result;
  double probability;
} rq_measurement;

/* types.rs:238-245 SimulationMetrics, plus device-side figures. */
typedef struct {
  uint32_t num_qubits;
  uint32_t _pad;
  uint64_t gate_count;         /* unitary gates of the input circuit */
  uint64_t execution_time_ns;  /* host wall clock, like simulator.rs:60,82 */
  uint64_t peak_memory_bytes;  /* QuantumState::estimate_memory */
  double gates_per_second;
  uint64_t gates_fused;        /* input gates folded into fused passes */
  uint64_t passes;             /* full-state kernel passes launched */
  double device_time_ms;       /* HIP-event time of the device work */
  double hbm_bytes;            /* algorithmic HBM bytes of the passes */
} rq_metrics;

/* types.rs:190-195 PauliOp and one (QubitIndex, PauliOp) factor of a
 * PauliString (types.rs:201-203). */
typedef enum { RQ_PAULI_I = 0, RQ_PAULI_X = 1, RQ_PAULI_Y = 2, RQ_PAULI_Z = 3 } rq_pauli_op;
typedef struct {
  uint32_t qubit;
  int32_t op;
} rq_pauli_factor;

typedef struct rq_state rq_state; /* opaque: owns the device amplitudes */

/* ---- errors / device -------------------------------------------------- */
const char* rq_last_error(void);
/* Number of qubits the device can hold for a precision (288 GB HBM3E:
 * 34 for c64, 35 for c32); the reference's fixed caps are MAX_QUBITS=32
 * (state.rs:15) and MAX_QUBITS_F32=33 (mixed_precision.rs:185). */
uint32_t rq_max_qubits(int precision);
int rq_device_synchronize(void);

/* ---- construction (state.rs:31-96, mixed_precision.rs:206-283) -------- */
/* QuantumState::new (with_seed=0) / new_with_seed (with_seed=1): |0..0>.
 * The seed itself is not needed (draws are inputs); with_seed only selects the
 * reference's validation (new(0) errors, new_with_seed(0) does not, f64). */
int rq_state_create(uint32_t num_qubits, int precision, int with_seed, rq_state** out);
/* QuantumState::from_amplitudes: len must equal 2^num_qubits. host_amps holds
 * rq_c64 (RQ_C64) or rq_c32 (RQ_C32) values. */
int rq_state_from_amplitudes(const void* host_amps, size_t len, uint32_t num_qubits, int precision,
                             rq_state** out);
void rq_state_destroy(rq_state* s);
uint32_t rq_state_num_qubits(const rq_state* s);      /* state.rs:102 */
int rq_state_precision(const rq_state* s);
size_t rq_estimate_memory(uint32_t num_qubits, int precision); /* state.rs:144-146 */
/* state_vector() (state.rs:110): device -> host copy of all amplitudes. */
int rq_state_read(const rq_state* s, void* host_amps);
/* amplitudes_mut() (state.rs:118) write-back: host -> device copy. */
int rq_state_write(rq_state* s, const void* host_amps);
/* Reset to |0..0> and clear the measurement record. */
int rq_state_set_zero(rq_state* s);
/* dst <- src (same size and precision). */
int rq_state_copy(const rq_state* src, rq_state* dst);
/* QuantumStateF32::from_f64 (mixed_precision.rs:260-278): a new complex-f32
 * state with each amplitude of the complex-f64 state src rounded to nearest
 * f32, src's measurement record and gate count 0.  Error: src not c64. */
int rq_state_from_f64(const rq_state* src, rq_state** out);
/* QuantumStateF32::to_f64 (mixed_precision.rs:280-288): a new complex-f64
 * state with each amplitude widened exactly (through from_amplitudes, so the
 * complex-f64 qubit limit applies) and no measurement record.  Error: src not
 * c32. */
int rq_state_to_f64(const rq_state* src, rq_state** out);
/* QuantumStateF32::precision_error_bound (mixed_precision.rs:325-327);
 * 0 for c64 states. */
double rq_precision_error_bound(const rq_state* s);

/* ---- gates (state.rs:158-276) ------------------------------------------ */
/* QuantumState::apply_gate.  Measure/Reset consume `draw`; when the gate is
 * a Measure, *out receives the outcome and *has_out is set to 1. */
int rq_apply_gate(rq_state* s, const rq_gate* g, double draw, rq_measurement* out, int* has_out);
/* apply_single_qubit_gate (state.rs:214-229); m is row-major [2][2]. */
int rq_apply_1q(rq_state* s, uint32_t qubit, const rq_c64* m);
/* apply_two_qubit_gate (state.rs:238-276); m is row-major [4][4]. */
int rq_apply_2q(rq_state* s, uint32_t q1, uint32_t q2, const rq_c64* m);

/* ---- circuits (simulator.rs:56-101) ------------------------------------ */
/* Simulator::run_with_config on an existing state: the state is first set to
 * |0..0>, then the gates run in order.  Measure/Reset gates consume draws in
 * order; measurements receive up to *nout_cap outcomes, *nout the count.
 * fusion=0: one exact kernel per gate (the reference's arithmetic, without
 * FMA contraction).  fusion=1: the fused schedule (gates folded into
 * LDS/register-resident tiles; rounding differs at the 1e-15 level).
 * fusion=2: dense 5-qubit blocks, each a 32x32 complex unitary applied on
 * the matrix cores (complex-f32 states of >= 10 qubits; others use 1). */
int rq_run_circuit(rq_state* s, const rq_gate* gates, size_t ngates, const double* draws, size_t ndraws,
                   rq_measurement* measurements, size_t nout_cap, size_t* nout, int fusion,
                   rq_metrics* metrics);
/* As rq_run_circuit but continue from the current state (no |0..0> reset). */
int rq_apply_circuit(rq_state* s, const rq_gate* gates, size_t ngates, const double* draws, size_t ndraws,
                     rq_measurement* measurements, size_t nout_cap, size_t* nout, int fusion,
                     rq_metrics* metrics);
/* types.rs:249-253 NoiseModel. */
typedef struct {
  double depolarizing_rate;
  double bit_flip_rate;
  double phase_flip_rate;
} rq_noise;
/* simulator.rs:56-101 with config.noise set: rq_run_circuit, and after every
 * gate the stochastic channel of apply_noise (simulator.rs:178-221) on each
 * qubit the gate touches: depolarising (one draw; below the rate a second
 * draw picks X / Y / Z at 1/3, 2/3), bit flip (one draw), phase flip (one
 * draw); a channel with rate 0 consumes no draw.  Measurement and noise draws
 * come from `draws` in the reference's order (the gate's own draw first).
 * The channel's Paulis are applied like gates but not counted in
 * metrics->gate_count; *draws_used receives the number of draws consumed. */
int rq_run_circuit_noisy(rq_state* s, const rq_gate* gates, size_t ngates, const rq_noise* noise,
                         const double* draws, size_t ndraws, size_t* draws_used, rq_measurement* measurements,
                         size_t nout_cap, size_t* nout, int fusion, rq_metrics* metrics);
/* circuit.rs:153-184 QuantumCircuit::depth. */
uint32_t rq_circuit_depth(uint32_t num_qubits, const rq_gate* gates, size_t ngates);

/* ---- StateCheckpoint (replay.rs:251-294) -------------------------------
 * The byte format of StateCheckpoint::capture: 2^n amplitudes as
 * interleaved (re, im) f64 pairs, little-endian, 16 bytes each (complex-f32
 * states are widened exactly).  rq_state_checkpoint writes 16 * 2^n bytes;
 * rq_state_from_checkpoint restores (replay.rs:270-287) into a new state of
 * the given precision (complex-f32 narrows each f64 to nearest);
 * nbytes must be 16 * 2^n, else InvalidStateVector. */
size_t rq_state_checkpoint_bytes(const rq_state* s);
int rq_state_checkpoint(const rq_state* s, void* out);
int rq_state_from_checkpoint(const void* bytes, size_t nbytes, int precision, rq_state** out);

/* ---- Grover search (crates/ruqu-algorithms/src/grover.rs) ---------------
 * rq_negate_amplitudes: the oracle step grover.rs:130-139 performs through
 * amplitudes_mut: a[i] = -a[i] for every listed index (an index listed twice
 * is negated twice); indices >= 2^n give InvalidStateVector.
 * rq_negate_amplitudes_from: a[i] = -a[i] for i >= start (grover.rs:146-155
 * with start = 1). */
int rq_negate_amplitudes(rq_state* s, const uint64_t* indices, size_t count);
int rq_negate_amplitudes_from(rq_state* s, uint64_t start);
/* grover.rs:69-74 optimal_iterations (max(round(pi/4/theta - 1/2), 1),
 * theta = asin(sqrt(M/N)); saturates like Rust's `as u32`). */
uint32_t rq_grover_optimal_iterations(uint32_t num_qubits, size_t num_targets);
/* grover.rs:42-55 GroverResult without the state. */
typedef struct {
  uint64_t measured_state;
  int32_t target_found;
  uint32_t num_iterations;
  double success_probability;
} rq_grover_result;
/* grover.rs:98-183 run_grover: iterations < 0 selects optimal_iterations;
 * draws[num_qubits] are the uniforms of the final measurement of qubits
 * 0..n-1 (grover.rs:187-200); a target >= 2^n is an argument error (an
 * assert in the reference); an empty target list with iterations < 0 is an
 * argument error (the reference's count is unbounded).  state_out (optional)
 * receives the final state (the caller destroys it). */
int rq_grover_run(uint32_t num_qubits, const uint64_t* targets, size_t num_targets, int64_t iterations,
                  int precision, int with_seed, const double* draws, rq_grover_result* out, rq_state** state_out);

/* ---- seeded RNG (state.rs:68: StdRng::seed_from_u64; rand 0.8.6,
 * rand_chacha 0.3.1, rand_core 0.6.4 per the reference's Cargo.lock) ------
 * The draws of a seeded QuantumState / Simulator: rq_stdrng_next_f64 gives
 * the stream of rng.gen::<f64>() (state.rs:302, simulator.rs:187-214). */
typedef struct rq_stdrng {
  uint32_t key[8];   /* ChaCha12 key (seed_from_u64's PCG32 expansion) */
  uint64_t block;    /* next block counter */
  uint32_t buf[64];  /* 4 blocks of output words */
  uint32_t index;    /* next word in buf (64: empty) */
} rq_stdrng;
void rq_stdrng_seed(rq_stdrng* r, uint64_t seed);
uint64_t rq_stdrng_next_u64(rq_stdrng* r);
uint32_t rq_stdrng_next_u32(rq_stdrng* r);
void rq_stdrng_next_f64(rq_stdrng* r, double* out, size_t n);
/* Simulator::run_shots' draws (simulator.rs:122-127): shot s uses a StdRng
 * seeded with base_seed + s (wrapping); out[s*per_shot + i] is its i-th f64. */
void rq_stdrng_shot_draws(uint64_t base_seed, size_t shots, size_t per_shot, double* out);
/* The ChaCha block function (16 input words, `rounds` rounds) under the RNG. */
void rq_chacha_block(const uint32_t* in, int rounds, uint32_t* out);

/* ---- measurement (state.rs:288-352) ------------------------------------ */
int rq_measure(rq_state* s, uint32_t qubit, double draw, rq_measurement* out);
/* state.rs:329-335: draws[q] is consumed by qubit q, qubit 0 first. */
int rq_measure_all(rq_state* s, const double* draws, rq_measurement* out);
int rq_reset_qubit(rq_state* s, uint32_t qubit, double draw);
/* Number of recorded measurements (state.rs:139-141) and a copy of them. */
size_t rq_measurement_record(const rq_state* s, rq_measurement* out, size_t cap);
/* Shots without re-simulation: each shot is measure_all (state.rs:329-335)
 * on a copy of the current state, i.e. Simulator::run_shots
 * (simulator.rs:107-164) for a circuit without Measure/Reset gates.
 * draws[shot*n + q] is the draw of qubit q; bits[shot*n + q] receives the
 * outcome.  *near_ties (optional) counts decisions whose draw lay within
 * 1e-12 of the threshold (where rounding order could flip the outcome).
 * The state itself is not collapsed. */
int rq_sample_shots(const rq_state* s, const double* draws, size_t shots, uint8_t* bits, uint64_t* near_ties);

/* ---- observables (state.rs:123-137, 362-435) --------------------------- */
int rq_probabilities(const rq_state* s, double* out);
int rq_probability_of_qubit(const rq_state* s, uint32_t qubit, double* out);
/* expectation_value of one PauliString (ops in order, duplicates allowed). */
int rq_expectation(const rq_state* s, const rq_pauli_factor* ops, size_t nops, double* out);
/* expectation_hamiltonian: term k = coeffs[k] * ops[offsets[k] .. offsets[k+1]). */
int rq_expectation_hamiltonian(const rq_state* s, const double* coeffs, const uint64_t* offsets,
                               const rq_pauli_factor* ops, size_t nterms, double* out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, code that forbids hipMalloc): the same result
 * as rq_expectation_hamiltonian, bit for bit, with its reduction partials in a
 * fixed device array of the library (32 x 2049 doubles, 512.5 KiB per GPU,
 * held until the process ends) instead of the state's grow-on-demand buffer.
 * Any state size and term count fits (terms go 32 per launch, partials are
 * capped at 2048 blocks).  Calls that overlap on one GPU share that array and
 * race; calls on one stream, or one after another, are safe. */
int rq_expectation_hamiltonian_noalloc(const rq_state* s, const double* coeffs, const uint64_t* offsets,
                                       const rq_pauli_factor* ops, size_t nterms, double* out);
/* <Z_q> for every qubit in one pass (out has num_qubits entries). */
int rq_expectation_z_all(const rq_state* s, double* out);
int rq_normalize(rq_state* s);                                  /* state.rs:415-423 */
int rq_fidelity(const rq_state* a, const rq_state* b, double* out); /* state.rs:426-435 */
int rq_norm_sq(const rq_state* s, double* out);

/* ---- QAOA MaxCut (crates/ruqu-algorithms/src/qaoa.rs) ------------------- */
/* qaoa.rs:34-40 Graph: num_nodes qubits, edges (ei[k], ej[k]) with weight w[k]. */
typedef struct {
  uint32_t num_nodes;
  uint32_t num_edges;
  const uint32_t* ei;
  const uint32_t* ej;
  const double* w;
} rq_graph;
typedef struct rq_qaoa rq_qaoa; /* one instance: graph + its device state */
/* qaoa.rs:96-110 QaoaResult (angles and history are separate outputs). */
typedef struct {
  double best_cut_value;
  uint64_t best_index;   /* best_bitstring[v] = (best_index >> v) & 1 */
  uint32_t iterations;   /* energy_history length */
  int32_t converged;
  uint64_t evaluations;  /* evaluate_qaoa_cost calls */
  double device_ms;      /* HIP-event time of all evaluations */
  double wall_ms;
} rq_qaoa_result;
/* Validates the graph like Graph::add_edge (qaoa.rs:56-60; node index out of
 * range -> RQ_ERR_INVALID_QUBIT instead of a panic) and allocates the state. */
int rq_qaoa_create(const rq_graph* g, int precision, rq_qaoa** out);
void rq_qaoa_destroy(rq_qaoa* q);
/* The state left by the last evaluation (owned by the instance). */
rq_state* rq_qaoa_state(rq_qaoa* q);
/* evaluate_qaoa_cost (qaoa.rs:179-201).  mode 0: the reference's arithmetic
 * (bit-identical f64 amplitudes); mode 1: fused mixer + one-pass energy. */
int rq_qaoa_cost(rq_qaoa* q, const double* gammas, const double* betas, uint32_t p, int mode, double* cost);
/* run_qaoa (qaoa.rs:217-308).  init_* may be null (0.5 each, qaoa.rs:221-222);
 * history has room for max_iter values. */
int rq_qaoa_run(rq_qaoa* q, uint32_t p, uint32_t max_iter, double lr, int mode, const double* init_gammas,
                const double* init_betas, double* gammas_out, double* betas_out, double* history,
                rq_qaoa_result* res);
/* argmax_i |a_i|^2 with ties to the larger index (qaoa.rs:249-255). */
int rq_argmax_probability(const rq_state* s, uint64_t* index, double* prob);
/* <prod_{q in masks[k]} Z_q> for each mask, one FP64 pass per 32 masks. */
int rq_expectation_zmasks(const rq_state* s, const uint64_t* masks, size_t n, double* out);

/* ---- batched QAOA / VQE (configs with thousands of small instances) ---- */
/* Many independent instances (2..24 qubits, QAOA MaxCut or VQE with the
 * Ry-CZ ansatz) in complex-f32 with FP64 reductions, evaluated by one
 * persistent kernel per optimizer round; optimizer steps follow run_qaoa
 * (qaoa.rs:217-308) and run_vqe (vqe.rs:148-216) per instance. */
typedef struct rq_batch rq_batch;
int rq_batch_create(rq_batch** out);
void rq_batch_destroy(rq_batch* b);
/* init_* may be null: 0.5 each (qaoa.rs:221-222). */
int rq_batch_add_qaoa(rq_batch* b, const rq_graph* g, uint32_t p, const double* init_gammas,
                      const double* init_betas, uint32_t* id);
/* Ry-CZ ansatz: per layer Ry(params[l*n+q]) on every qubit, then CZ(q, q+1);
 * Hamiltonian term k = coeffs[k] * ops[offsets[k] .. offsets[k+1]).
 * init_params may be null: 0.1 each (vqe.rs:152). */
int rq_batch_add_vqe(rq_batch* b, uint32_t n, uint32_t layers, const double* coeffs, const uint64_t* offsets,
                     const rq_pauli_factor* ops, size_t nterms, const double* init_params, uint32_t* id);
/* Energy (and QAOA argmax index) of every instance at its current parameters. */
int rq_batch_evaluate(rq_batch* b, double* energies, uint64_t* argmax);
/* max_iter optimizer iterations for every instance; use_graph = 1 captures one
 * iteration as a hipGraph and replays it. */
int rq_batch_optimize(rq_batch* b, uint32_t max_iter, double qaoa_lr, double vqe_lr, double vqe_threshold,
                      int use_graph, double* device_ms);
/* Results of one instance: best energy (QAOA max / VQE min), QAOA best
 * bitstring index, history length, converged flag, final parameters, the
 * parameters of the best energy, energy history. */
int rq_batch_result(rq_batch* b, uint32_t id, double* best, uint64_t* best_index, uint32_t* iterations,
                    int* converged, double* params, double* best_params, double* history);
/* rq_batch_result for every instance at once; params / best_params hold all
 * instances' parameters back to back in instance order. */
int rq_batch_results(rq_batch* b, double* best, uint64_t* best_index, uint32_t* iterations, int* converged,
                     double* params, double* best_params);
int rq_batch_set_params(rq_batch* b, uint32_t id, const double* params);
/* Every instance's parameters at once: params holds them back to back in
 * instance order (count = the batch's total parameter count); one upload. */
int rq_batch_set_all_params(rq_batch* b, const double* params, size_t count);
uint32_t rq_batch_size(const rq_batch* b);
/* Exact MaxCut value of every QAOA instance (device brute force over all 2^n
 * bitstrings, FP64 weights; 0 for VQE instances): the denominator of
 * approximation ratios (no reference counterpart). */
int rq_batch_maxcut(rq_batch* b, double* out);
/* HIP-event time of the last evaluate/optimize call. */
double rq_batch_last_ms(const rq_batch* b);
/* LDS tile of the states above 13 qubits: 2^tq amplitudes per workgroup,
 * ceil(n_max/2) <= tq <= 13 (no reference counterpart; the autotuner's knob).
 * A state of n qubits whose second tiling would read runs of fewer than 16
 * amplitudes (2 tq - n < 4) gets a wider tile, up to 13 qubits. */
int rq_batch_set_tile(rq_batch* b, int tq);
/* The same with every state above 13 qubits on tiles of exactly tq qubits. */
int rq_batch_set_tile_exact(rq_batch* b, int tq);
int rq_batch_get_tile(const rq_batch* b);
/* Host-only inspection (no device call; no reference counterpart): the launch
 * plan of one round (round < 0: the plain evaluation; else round r of an
 * optimizer iteration) as the device receives it.  info[8] = {nsmall, nwide,
 * ntiles, nlaunches, nrounds, nouts, tq, nvtiles}; info64[2] = {amplitudes of
 * the round's wide states, cut-table floats}; small/wide items are 16-byte
 * records; wtile0/wstate: first tile and state offset per wide item; wmap:
 * wide item of each tile; pass_grid: tiles of each layer-pass launch (one per
 * pass and tile width); vmap: the round's tiles of VQE items.  Null outputs
 * are skipped (call once for the counts). */
int rq_batch_plan_export(rq_batch* b, int round, int32_t* info, int64_t* info64, void* small_items,
                         void* wide_items, int32_t* wtile0, int64_t* wstate, int32_t* wmap, int32_t* pass_grid,
                         int32_t* vmap);
/* Host-only: instance (48 B), edge (16 B), Pauli-term (32 B) and parameter
 * records in device layout; counts[4] = {instances, edges, terms, params}. */
int rq_batch_data_export(rq_batch* b, void* insts, void* edges, void* terms, double* params, size_t* counts);
/* Host-only: the pass structure of a round's wide items (states above 13
 * qubits): first / last pass, the state the first pass reads (a shared prefix
 * state when the first pass is > 0), and every launch's tile list
 * (concatenated; launch i at pass_off[i]). */
int rq_batch_plan_passes(rq_batch* b, int round, int32_t* wfirst, int32_t* wlast, int64_t* wsrc, int32_t* ptile,
                         int32_t* pass_off);
/* Host-only: the pass and the tile qubits of each layer-pass launch of a
 * round (nlaunches entries each). */
int rq_batch_plan_launches(rq_batch* b, int round, int32_t* pass, int32_t* tile_qubits);

/* ---- fused schedule (no reference counterpart: the reference's optimizer
 * only folds adjacent 1q gates, optimizer.rs:46-94, and is not called by
 * Simulator) ------------------------------------------------------------ */
/* Tile size m (qubits per LDS/register tile, 4..12) and count c of lowest
 * qubits kept in every tile; applies to later rq_run_circuit(fusion=1). */
int rq_fusion_configure(int m, int c);
/* Host-only: schedule size of gates[0..n) (no Measure/Reset). */
int rq_fusion_plan_info(uint32_t num_qubits, const rq_gate* gates, size_t ngates, int m, int c, uint64_t* stages,
                        uint64_t* phases, uint64_t* ops);
/* Host-only: executes the program the fused kernel would receive (packed op
 * records, folded gates, matrix-core phases at the current threshold, tuning
 * key 4) on host amplitudes with the kernel's data mapping and lane layout
 * (planner and packing check without a GPU). */
int rq_fusion_emulate(uint32_t num_qubits, int precision, const rq_gate* gates, size_t ngates, int m, int c,
                      void* host_amps);
/* Host-only: size of that program: op records, matrix-core phases and the
 * composed unitaries they select from. */
int rq_fusion_packed_info(uint32_t num_qubits, int precision, const rq_gate* gates, size_t ngates, int m, int c,
                          uint64_t* records, uint64_t* mf_phases, uint64_t* mf_unitaries);
/* Host-only: that program's records by kind (counts[9]): 0 dense 2x2, 1 dense
 * 4x4, 2 one-bit diagonal, 3 two-bit diagonal, 4 CNOT, 5 CZ, 6 padding,
 * 7 factored 2x2 (unit-diagonal or unit-antidiagonal form), 8 pending-diagonal
 * tables over the register bits. */
int rq_fusion_packed_kinds(uint32_t num_qubits, int precision, const rq_gate* gates, size_t ngates, int m, int c,
                           uint64_t* counts);
/* Host-only: that program's records: the first 32-bit word of each (kind:3 |
 * t0:2 | t1:2 | cloc:2 | cpos:5 | dloc0:2 | dpos0:6 | dloc1:2 | dpos1:6;
 * locations 0 none, 1 register bit, 2 thread bit, 3 tile-id bit) and the
 * phase each belongs to (-1: a matrix-core phase's record).  count receives
 * the number of records; RQ_ERR_ARGUMENT when it exceeds cap. */
int rq_fusion_packed_records(uint32_t num_qubits, int precision, const rq_gate* gates, size_t ngates, int m, int c,
                             uint32_t* words, int32_t* phase, size_t cap, size_t* count);

/* Process-wide tuning knobs (no reference counterpart; set by the autotuner):
 * key 0 = streaming kernels' amplitude groups per thread (1, 2, 4, 8),
 * key 1 = non-temporal HBM access (0/1), key 2 = fused tile qubits m,
 * key 3 = fused low qubits c, key 4 = fused matrix-core threshold: complex-f32
 * phases whose ops cost at least this many VALU instructions per wave run as
 * one composed 16x16 unitary on the matrix cores (-1: never), key 5 = tile-qubit
 * seeds the fused planner tries per stage (0..4096, default 512; 0: plain
 * greedy scan; the plan stays valid for any value, only the stage count
 * changes), key 6 = workgroup order of the two-stream exact 1q kernels (-1:
 * the measured per-distance chunk table, 0: linear, a power of two: each XCD
 * takes that many consecutive blocks at a time).
 * rq_tune_get returns -2 for an unknown key. */
int rq_tune_set(int key, int value);
int rq_tune_get(int key);

/* ---- timing helpers for the measurement harness ------------------------ */
/* Time `reps` back-to-back launches of the exact 1q kernel on `qubit` with
 * matrix m (HIP events on the state's stream); *avg_ms = mean per launch. */
int rq_time_1q(rq_state* s, uint32_t qubit, const rq_c64* m, int reps, double* avg_ms);
/* Time `reps` launches of a plain device copy of the state's bytes (read N,
 * write N) as the achievable-bandwidth reference; *avg_ms per launch. */
int rq_time_copy(rq_state* s, int reps, double* avg_ms);

/* ---- learned warm start (no reference counterpart) ---------------------
 * A graph network predicting QAOA angles (gamma_1..8, beta_1..8) from a
 * weighted graph and the depth p: 3 message-passing layers of width 256 with
 * bfloat16 matrix-core products (FP32 accumulation), trained on the device
 * by its own backward pass and Adam.  Its specification (parameter layout,
 * features, loss) is the FP64 restatement in oracle/mpnn_oracle.py. */
typedef struct rq_mpnn rq_mpnn;
typedef struct {
  int32_t ngraphs, nnodes;
  int64_t nedges;
  const int32_t* node_off; /* ngraphs+1: nodes of graph g are node_off[g] .. node_off[g+1]-1 */
  const int32_t* row_ptr;  /* nnodes+1: incoming edges of node v (CSR by destination) */
  const int32_t* col;      /* nedges: source node of each edge (both directions listed) */
  const float* w;          /* nedges: edge weight */
  const float* x;          /* nnodes x 32 node features (degree/16, weighted degree/8, 0...) */
  const float* gfeat;      /* ngraphs x 32 graph features (onehot(p), n/22, edge density, 0...) */
  const int32_t* p;        /* ngraphs: QAOA depth, 1..8 */
} rq_mpnn_batch;
size_t rq_mpnn_num_params(void);
int rq_mpnn_create(rq_mpnn** out);
void rq_mpnn_destroy(rq_mpnn* n);
/* Flat FP32 parameters (layout of oracle/mpnn_oracle.py); set also resets Adam. */
int rq_mpnn_set_params(rq_mpnn* n, const float* params);
int rq_mpnn_get_params(const rq_mpnn* n, float* params);
/* angles: ngraphs x 16 (gammas then betas; entries l >= p are unused). */
int rq_mpnn_forward(rq_mpnn* n, const rq_mpnn_batch* b, float* angles);
/* Masked-MSE loss against targets (ngraphs x 16) and its gradient (FP32,
 * rq_mpnn_num_params() values; grad may be null). */
int rq_mpnn_gradient(rq_mpnn* n, const rq_mpnn_batch* b, const float* targets, float* loss, float* grad);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the batch; loss before the step. */
int rq_mpnn_train_step(rq_mpnn* n, const rq_mpnn_batch* b, const float* targets, float lr, float* loss);

#ifdef __cplusplus
}
#endif

#endif /* RUQU_GPU_H */
