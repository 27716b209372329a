// Fused execution of a circuit: gates are folded into "stages".  A stage
// covers a set S of m tile qubits (always containing the lowest `c` qubits so
// that every tile is made of contiguous runs in HBM).  One workgroup loads one
// tile (the 2^m amplitudes that share the values of the other qubits) into
// registers, applies every gate of the stage, and writes the tile back: one
// HBM read + write per stage instead of one per gate.
//
// Inside a tile each thread holds R = 16 amplitudes whose index bits are the
// 4 "register qubits" of the current phase; gates act on register qubits in
// registers, and a phase change re-maps the tile through LDS.  Diagonal gates
// (Rz, Phase, S, T, Z, CZ, Rzz) and the controls of CNOT never need to be in
// S: their factor only reads index bits, which are known for every amplitude
// (register bit, thread bit, or a per-tile constant).
#pragma once

#include <cstdint>
#include <vector>

#include "internal.h"

namespace rq {

constexpr int FUSE_R_BITS = 4;         // register qubits per phase
constexpr int FUSE_R = 1 << FUSE_R_BITS;
constexpr int FUSE_MAX_M = 12;         // largest tile (qubits): 256 threads x 16 amplitudes
constexpr int FUSE_MAX_OUTER = 48;

// U1: 2x2 on a register qubit (optionally controlled); U2: 4x4 on two register
// qubits; D1/D2: diagonal factor read from one/two index bits; CX: controlled
// swap of a register-qubit pair (CNOT); CZ: sign flip where both bits are 1.
enum FOpKind : int32_t { FOP_U1 = 0, FOP_U2 = 1, FOP_D1 = 2, FOP_D2 = 3, FOP_CX = 4, FOP_CZ = 5 };

// Where an index bit of an amplitude comes from inside a phase.
enum BitLoc : int8_t { LOC_REG = 0, LOC_THR = 1, LOC_OUT = 2 };

struct FOp {
  int32_t kind;
  int32_t t0, t1;      // register bits of the targets (U1: t0; U2: t0 = q1, t1 = q2)
  int32_t mat;         // offset of the matrix / diagonal in the pool (complex entries)
  uint32_t ctl_reg;    // controls that are register bits (mask over k)
  uint32_t ctl_thr;    // controls that are thread bits (mask over thread id)
  uint64_t ctl_out;    // controls that are outer qubits (mask over the global index)
  int8_t dloc0, dloc1; // diagonal bits: location kind
  int8_t pad0, pad1;
  int32_t dpos0, dpos1;  // bit index within that location
};

// One phase of a stage.  128 bytes, the fields the kernel reads first, so
// that a workgroup fetches a phase with two 64-byte scalar loads.
struct FPhase {
  int32_t op_begin, op_end;
  int32_t direct;                // 1: as first / last phase, access HBM in this phase's mapping
  int32_t mf_off;                // -1: ops on the VALU; else pool offset of the phase's matrix-core operands
  uint32_t mf_sel;               // matrix-core phase: the uniform bits that select its matrix (3 x 8 bits)
  int32_t regt[FUSE_R_BITS];     // tile bits holding register bits 0..3
  int32_t greg[FUSE_R_BITS];     // global qubits of register bits 0..3 (stage S of regt, set at packing)
  int32_t thrt[FUSE_MAX_M - FUSE_R_BITS];  // tile bits holding thread-id bits 0..m-5 (0 past m-5)
  int32_t gthr[FUSE_MAX_M - FUSE_R_BITS];  // global qubits of thread-id bits 0..m-5 (S of thrt)
  int32_t pad[3];
  // LDS byte offset of register index k's part of the tile index (the
  // swizzled address of the register bits set in k), filled at upload for
  // the state's element size
  uint32_t lreg[1 << FUSE_R_BITS];
};
static_assert(sizeof(FPhase) == 192, "phase record: three 64-byte scalar loads");

struct FStage {
  int32_t m;            // tile qubits
  int32_t nphase;
  int32_t phase0;       // first phase in the phase array
  int32_t nouter;
  int32_t S[FUSE_MAX_M];          // global qubit of each tile bit
  int32_t outer[FUSE_MAX_OUTER];  // global qubit of each outer bit (tile id bits)
};

struct FusionStats {
  uint64_t unitary = 0;  // unitary input gates
  uint64_t fused = 0;    // input gates folded (unitary - passes when positive)
  uint64_t passes = 0;   // stage launches (full-state HBM passes)
  double hbm_bytes = 0;  // algorithmic bytes: 2*N*s per stage
};

// Tunables of the fused schedule.
struct FusionConfig {
  int m = 12;      // tile qubits
  int c = 5;       // lowest qubits always in the tile
  int mf = 768;    // complex-f32 phases whose ops cost at least this many VALU instructions per wave run on
                   // the matrix cores as one composed 16x16 unitary; -1: never
  int seeds = 512; // random tile-qubit seeds tried per stage (0: the plain greedy scan only)
};

// Executes gates[0..ngates) with the fused schedule; Measure/Reset gates run
// exactly (via measure/reset on the state) between fused segments.
int run_fused(State& s, const rq_gate* gates, size_t ngates, const double* draws, size_t ndraws, size_t* cursor,
              rq_measurement* meas, size_t cap, size_t* produced, FusionStats* st,
              bool zero_first = false);
// Same contract with dense 5-qubit blocks on the matrix cores (dense.cpp).
int run_dense(State& s, const rq_gate* gates, size_t ngates, const double* draws, size_t ndraws, size_t* cursor,
              rq_measurement* meas, size_t cap, size_t* produced, FusionStats* st);

// Host-side plan of one fused segment (exposed for tests and the autotuner).
struct FusedPlan {
  std::vector<FStage> stages;
  std::vector<FPhase> phases;
  std::vector<FOp> ops;
  std::vector<rq_c64> pool;
  uint64_t unitary = 0;
};
int plan_fused(uint32_t nq, const rq_gate* gates, size_t ngates, const FusionConfig& cfg, FusedPlan* plan);

// Device launch of a planned segment (kernels_fused.hip).
int launch_fused_plan(State& s, const FusedPlan& plan);
// Plan (or reuse the state's cached program for the same gates and knobs),
// upload and launch one segment without Measure/Reset gates.
// With *zero_in set, the state is taken as |0..0> without reading it (its
// memory may hold anything): on a complex-f32 state the first stage makes the
// amplitudes, on a complex-f64 state a clearing pass runs first; *zero_in is
// cleared once a stage ran (a segment without stages leaves it set).
int run_fused_segment(State& s, const rq_gate* gates, size_t ngates, uint64_t* unitary, uint64_t* stages,
                      bool* zero_in = nullptr);

// Host execution of the program the kernel receives for a plan (packed op
// records, folded gates, matrix-core phases) with the kernel's data mapping
// and lane layout; `amps` holds rq_c64 or rq_c32 values per precision.  For
// checking the packing on machines without a GPU.
int emulate_packed(const FusedPlan& plan, int precision, int mf_min, void* amps);
// Size of the packed program: op records, matrix-core phases and their
// composed unitaries.
// Records of the packed program by kind: counts[0..7] per record kind (2x2,
// 4x4, 1-bit diagonal, 2-bit diagonal, CNOT, CZ, padding, factored 2x2),
// counts[8] the pending-diagonal tables (counted apart from kind 2).
int packed_kinds(const FusedPlan& plan, int precision, int mf_min, uint64_t* counts);
int packed_records(const FusedPlan& plan, int precision, int mf_min, uint32_t* words, int32_t* phase, size_t cap,
                   size_t* count);
int packed_info(const FusedPlan& plan, int precision, int mf_min, uint64_t* records, uint64_t* mf_phases,
                uint64_t* mf_unitaries);

// Measurement helpers shared with the fused runner (state.cpp).
int measure_state(State& s, uint32_t q, double draw, rq_measurement* out);
int reset_state(State& s, uint32_t q, double draw);
FusionConfig& fusion_config();

}  // namespace rq
