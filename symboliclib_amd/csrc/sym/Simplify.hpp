// Algebraic normalisation of kernel statements (the job of the reference's src/Simplify.cpp).
// The result fixes the order of the floating-point operations the emitters write out.
#pragma once

#include "Symbolic.hpp"

namespace Sym {

// Rebuild bottom-up: operations whose operands are all constants are evaluated, constants in
// sums and products are combined (a product with a zero factor is zero), and an expression
// whose algebraic hash is a small integer k is the constant k.  Algebraically equal
// sub-expressions come back as one node.
Symbolic foldConstants(const Symbolic& x);

// foldConstants, and at every sum and product a normal form:
//  * sums: terms of nested sums, negations and constant multiples are collected as
//    (coefficient, term) pairs in order of first appearance; algebraically equal terms are
//    merged and zero terms dropped; while a factor is shared by two or more terms, the one
//    shared by the most (then the most complex) is factored out; terms with a negative
//    coefficient are subtracted as one sum;
//  * products and quotients: the powers of the factors of nested products, quotients, square
//    roots and constant powers are added up per base; integer powers become repeated factors,
//    half powers one square root, negative powers the denominator.
Symbolic simplify(const Symbolic& x);

}  // namespace Sym
