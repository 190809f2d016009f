#pragma once

#include <vector>

#include "Symbolic.hpp"

namespace Sym {

// Split a set of expressions into blocks: every sub-expression that is reached through two or
// more parents (by node identity) and whose complexity reaches `complexityThreshold` becomes
// its own block `ASSIGN(VAR(id, variableGroupId), expr)`; FIXED blocks stay indivisible.
// Follows src/Decomposition.cpp:234-490 of the reference (count, children, merge/levels,
// extraction).  threshold 255 disables decomposition.
std::vector<ExpressionBlock> decompose(const std::vector<Symbolic>& expr, const unsigned char complexityThreshold = 3,
                                       const unsigned char variableGroupId = 255, const bool mergeSingleParent = true);

}  // namespace Sym
