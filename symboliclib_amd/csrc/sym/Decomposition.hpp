// Decomposition of a set of expressions into units of work (the job of the reference's
// src/Decomposition.cpp:234-490): an expression reached more than once whose complexity is at
// least a threshold is computed once, as a block assigning a fresh variable that its readers
// use instead; FIXED sub-expressions (sealed element blocks) are blocks of their own and never
// split.  Each block carries a shape hash (equal for blocks that differ only in the values of
// their leaves, so they can become instances of one kernel) and a height in the block graph.
#pragma once

#include <vector>

#include "Symbolic.hpp"

namespace Sym {

struct DecomposedBlock {
    // ASSIGN(VAR(id, group), expr), ASSIGN to a caller's output, a sealed BLOCK or an INSTANCE
    Symbolic code;
    // structure hash: leaves hash by kind only (variable or constant) and the operands of
    // commutative operations are ordered by their own shape hash, so blocks of equal shape
    // have equal trees position by position
    hash_t shape = 0;
    // 0 when the block reads no other block, else one more than the highest block it reads:
    // blocks of equal height never read each other
    int height = 0;
    std::vector<int> reads;  // indices of the blocks whose variables this block reads
};

// roots: the expressions to compute.  A sub-expression that occurs at least twice (counted
// once per distinct occurrence of its parent) with complexity >= threshold becomes a block of
// variable group `group`, unless its own cost (its operations outside other blocks) is below
// the threshold or, with inlineSingleReader, only one block reads it: those are written into
// their readers.  threshold 255: every root is one block, nothing is shared.
std::vector<DecomposedBlock> decompose(const std::vector<Symbolic>& roots, unsigned char threshold, int group,
                                       bool inlineSingleReader);

}  // namespace Sym
