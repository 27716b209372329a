// QAOA MaxCut on the device: kernel launchers (kernels_qaoa.hip) shared by
// the single-instance driver (qaoa.cpp).
#pragma once

#include <cstring>

#include "internal.h"

namespace rq {

// |+>^n: every amplitude set to c (the product h^n of the reference's H gates).
int qaoa_fill_uniform(State& s, double c);
// One phase-separator layer: a_i *= prod_e (bits differ ? ep_e : en_e), edge order.
int qaoa_cost_phase(State& s, const void* d_edges, int ne);
size_t qaoa_edge_record_bytes(int prec);
void qaoa_pack_edges(int prec, const uint32_t* ei, const uint32_t* ej, const rq_c64* en, const rq_c64* ep, int ne,
                     void* out);
size_t qaoa_weight_record_bytes();
void qaoa_pack_weights(const uint32_t* ei, const uint32_t* ej, const double* w, int ne, void* out);
// sum_i |a_i|^2 C(i) (FP64).
int qaoa_energy(State& s, const void* d_wedges, int ne, double* out);
// <prod_{q in mask} Z_q> for each mask (FP64, 32 masks per pass).
int reduce_zmasks(State& s, const uint64_t* masks, int nmasks, double* out);
// argmax_i |a_i|^2, ties to the larger index.
int argmax_probability(State& s, uint64_t* index, double* prob);

constexpr int QAOA_MAX_EDGES = 1024;

}  // namespace rq
