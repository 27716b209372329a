// Batched QAOA / VQE: many uneven instances evaluated together
// (kernels_batch.hip: a persistent kernel for LDS-resident states, per-tile
// launches for larger ones), optimised by device-side update kernels
// (batch.cpp), one optimizer iteration per hipGraph replay.
#pragma once

#include <cstdint>

namespace rq {

constexpr int BATCH_LDS_QUBITS = 13;       // states up to 2^13 c32 amplitudes live in LDS (64 KiB)
constexpr int BATCH_TILE_QUBITS = 12;      // default LDS tile of the HBM-resident ("wide") states
constexpr int BATCH_MIN_TILE_QUBITS = 9;
constexpr int BATCH_MAX_TILE_QUBITS = 13;
constexpr int BATCH_MAX_QUBITS = 24;       // two tile passes per layer need n <= 2 * tile qubits
constexpr int BATCH_THREADS = 256;
constexpr int BATCH_CUT_CHUNK = 4096;      // cut-table entries per workgroup of k_cut_tables

enum BatchKind : int32_t { BK_QAOA = 0, BK_VQE = 1 };

// Tile qubits of a wide state of n qubits when the batch's tile is `base`:
// the tiling B pass reads runs of 2^(2 tq - n) amplitudes, so the tile widens
// (up to BATCH_MAX_TILE_QUBITS) until a run fills a 128-byte line (16
// complex-f32 amplitudes, 2 tq - n >= 4).
inline int batch_item_tq(int n, int base) {
  const int need = (n + 5) / 2;
  return need > base ? (need < BATCH_MAX_TILE_QUBITS ? need : BATCH_MAX_TILE_QUBITS) : base;
}

struct BInst {
  int32_t kind;
  int32_t n;
  int32_t p;          // QAOA depth or VQE layers
  int32_t nparams;    // QAOA 2p (gammas then betas), VQE n*layers
  int32_t nterms;     // QAOA edges or VQE Pauli terms
  int32_t term_off;   // first edge / term
  int32_t param_off;  // first parameter (doubles)
  int32_t tq;         // n > BATCH_LDS_QUBITS: tile qubits of the state's passes (set when planning), else 0
  int64_t ctab_off;   // QAOA, n > BATCH_LDS_QUBITS: offset of the cut table (floats), else -1
  double wsum;        // QAOA: total edge weight
};

struct BEdge {
  uint32_t i, j;
  double w;
};

// Per-pass launch data of k_wide_pass: the pass's tiles (round tile ids, item
// by item) and every item's first / last pass and the state its first pass
// reads (amplitude offset; a shared prefix when the first pass is > 0).
struct WidePasses {
  const int32_t* ptile;
  const int32_t* first;
  const int32_t* last;
  const int64_t* src;
};

// One workgroup of k_cut_tables: entries [chunk * BATCH_CUT_CHUNK, ...) of
// instance inst's cut table.
struct BCutChunk {
  int32_t inst;
  uint32_t chunk;
};

struct BTerm {
  uint64_t x, zy;
  uint32_t ny, pad;
  double c;
};

// One circuit evaluation: instance, variant (0 base, 1 +pi/2, 2 -pi/2 on
// parameter `param`), and the slot of its energy in the output array.
struct BItem {
  int32_t inst;
  int32_t variant;
  int32_t param;
  int32_t out;
};

// Per-tile partial of a wide item: energy contribution, the tile's largest
// |a|^2 and its index.
struct BPart {
  double e;
  double pm;
  uint32_t idx;
  uint32_t pad;
};

}  // namespace rq
