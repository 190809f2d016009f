// Operation codes and per-operation properties of the symbolic IR.
//
// The numbering, arity, infix/commutative flags, printed symbols and decomposition costs
// follow src/Operations.hpp:9-85 of the reference, because the costs decide which
// sub-expressions become shared intermediates and the codes are part of the public API
// (`Symbolic(ADD, a, b)`).  The per-operation hash constants are derived here from the
// operation names.  INSTANCE is new: a lazily expanded instance of an expression template
// (one element's block of outputs), see Template.hpp.
#pragma once

#include <string>
#include <vector>

#include "Hashing.hpp"

namespace Sym {

typedef unsigned char OpType;

constexpr OpType NOOP = 0;
constexpr OpType CONST = 1;
constexpr OpType VAR = 2;
constexpr OpType ADD = 3;
constexpr OpType MUL = 4;
constexpr OpType MULC = 5;
constexpr OpType DIV = 6;
constexpr OpType SQRT = 7;
constexpr OpType SIN = 8;
constexpr OpType COS = 9;
constexpr OpType TAN = 10;
constexpr OpType COT = 11;
constexpr OpType ASIN = 12;
constexpr OpType ACOS = 13;
constexpr OpType ATAN = 14;
constexpr OpType EXP = 15;
constexpr OpType LOG = 16;
constexpr OpType POW = 17;
constexpr OpType ASSIGN = 18;
constexpr OpType GROUP = 19;
constexpr OpType CONDPOS = 20;
constexpr OpType CONDNONZERO = 21;
constexpr OpType CONDZERO = 22;
constexpr OpType NEG = 23;
constexpr OpType DEFINE = 24;
constexpr OpType BLOCK = 25;
constexpr OpType RECIPROCAL = 26;
constexpr OpType FIXED = 27;
constexpr OpType FUNCTIONHANDLE = 28;
constexpr OpType CALL = 29;
constexpr OpType INSTANCE = 30;
constexpr int NUM_OPS = 31;

struct OpInfo {
    unsigned char operands;  // 0 (leaf), 1, 2, 3 or 255 (any number)
    bool infix;
    bool commutative;
    bool leaf;
    std::string opSymbol;
    std::string code;
    hash_t hash;
    unsigned char cost;
};

extern const OpInfo OpInfos[NUM_OPS];

class Symbolic;

// Numeric value of operation `id` applied to already evaluated children.
double evaluate(const OpType id, const std::vector<double>& childValues);

/* Multiplication ************************************************************/
Symbolic operator*(const Symbolic& a, const Symbolic& b);
Symbolic operator*=(Symbolic& a, const Symbolic& b);
/* Addition ******************************************************************/
Symbolic operator+(const Symbolic& a, const Symbolic& b);
Symbolic operator+=(Symbolic& a, const Symbolic& b);
/* Division ******************************************************************/
Symbolic operator/(const Symbolic& a, const Symbolic& b);
Symbolic operator/=(Symbolic& a, const Symbolic& b);
/* Subtraction ***************************************************************/
Symbolic operator-(const Symbolic& a, const Symbolic& b);
Symbolic operator-=(Symbolic& a, const Symbolic& b);
Symbolic operator-(const Symbolic& a);
/* unary functions ***********************************************************/
Symbolic sqrt(const Symbolic& a);

}  // namespace Sym
