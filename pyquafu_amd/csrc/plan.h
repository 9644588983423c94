// Gate-run execution: turns a list of internal gates into kernel launches.
#pragma once

#include <vector>

#include "common.h"

namespace qfvm {

// Applies `gates` in order to the state (asynchronously on s.stream).
void apply_gates(State& s, const std::vector<Gate>& gates);
// apply_gates for the gate runs of qfvm_apply / qfvm_run: with option "dense_blocks", deep runs on
// few qubits become one dense gate each first. (The adjoint's sweeps, which graph mode captures,
// apply their gates as they are.)
void apply_circuit(State& s, const std::vector<Gate>& gates);
// The reverse sweep: applies `gates` in order to both s.psi and `lam`; a G_GRAD marker writes
// grads_dev[slot] = Im <lam| P (controls projected) |psi> for the two states at its position.
void apply_gates2(State& s, const std::vector<Gate>& gates, void* lam, double* grads_dev);

}  // namespace qfvm
