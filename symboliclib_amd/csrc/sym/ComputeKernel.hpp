// Kernel templates: one expression evaluated for many instances (the job of the reference's
// src/ComputeKernel.cpp:27-484).
//
// The instances of a group (blocks of equal shape) are walked position by position.  A
// position where every instance holds the same constant stays a literal; any other leaf
// becomes a per-instance column the template reads through a leaf variable:
//   VAR(k, -2)  output column k    (the variables an instance assigns)
//   VAR(k, -3)  input leaf k       (a variable an instance reads: column and offset in
//                                   inputLeaves)
//   VAR(k, -4)  constant column k  (constants that differ between instances)
// Positions whose instance nodes are the same in every instance (equal per-instance
// algebraic hashes) share one template node.  An input column equal to another stored column
// plus a constant (one input group in every instance) is not stored again: its leaf reads the
// stored column with that offset.  Tables are structure of arrays (column c, instance i at
// c * n + i).  The template then goes through optimizeTemplate.
#pragma once

#include <array>
#include <vector>

#include "Symbolic.hpp"

namespace Sym {

struct ColumnRef {
    int column;  // stored input column
    int offset;  // constant added to the stored variable index
};

struct KernelTemplate {
    size_t instances = 0;
    Symbolic code;  // BLOCK of statements (optimizeTemplate)
    std::vector<std::vector<std::array<int, 2>>> outputColumns;  // per output leaf: variable per instance
    std::vector<std::vector<std::array<int, 2>>> inputColumns;   // stored input columns
    std::vector<ColumnRef> inputLeaves;                          // per input leaf of the template
    std::vector<std::vector<double>> constantColumns;            // per constant leaf: value per instance
};

// instances: blocks of one group (equal shape); throws std::logic_error when they differ.
KernelTemplate buildKernelTemplate(const std::vector<Symbolic>& instances, bool simplifyTemplate);

// Template optimisation shared by the kernel builder and the direct assembly builders (the job
// of src/ComputeKernel.cpp:270-323): expand an INSTANCE, fold constants (algebraically equal
// sub-expressions become one node), turn every sub-expression used twice into a local
// definition (LOCAL_VAR, DEFINE), simplify each statement, definitions before their uses.
// Returns a BLOCK.
Symbolic optimizeTemplate(Symbolic templateExpression, bool simplifyTemplate);

}  // namespace Sym
