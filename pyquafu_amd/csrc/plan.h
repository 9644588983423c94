// Gate-run execution: turns a list of internal gates into kernel launches.
#pragma once

#include <vector>

#include "common.h"

namespace qfvm {

// Applies `gates` in order to the state (asynchronously on s.stream).
void apply_gates(State& s, const std::vector<Gate>& gates);

}  // namespace qfvm
