// Algebraic clean-up of expression templates (src/Simplify.hpp, src/Simplify.cpp:14-505 of the
// reference): flattening of sums and products, constant folding, merging of like terms,
// factoring out shared factors, power normalisation and negation extraction.  The result
// fixes the order of floating-point operations in the emitted kernels.
#pragma once

#include <vector>

#include "Symbolic.hpp"

namespace Sym {

Symbolic makeMul(const std::vector<Symbolic>& childs);
Symbolic makeAddition(std::vector<Symbolic> operands);
Symbolic makeSymbolic(const unsigned char op, const std::vector<Symbolic>& operands);

Symbolic flattenAdditions(const Symbolic& x);
Symbolic flattenMultiplications(const Symbolic& expr);
Symbolic flattenDivAndMul(const Symbolic& expr);
Symbolic optimizeProducts(const Symbolic& x);
Symbolic optimizeSum(const Symbolic& x);
Symbolic pullFactor(Symbolic x);
Symbolic findNegative(const Symbolic& expr);
Symbolic simplify(const Symbolic& x);

// Rebuild with constant folding; algebraically equal sub-expressions become one shared node.
Symbolic removeConstantExpressions(const Symbolic& x);
Symbolic referenceRedundant(const Symbolic& x);

}  // namespace Sym
