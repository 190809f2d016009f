// Direct assembly of element-sum Hessians into a Program (the scalable front-end).
//
// The drop-in path builds one expression per output nonzero: hessianSparse -> toSparseMatrix
// (src/SymbolicDifferentiation.hpp:234-278, src/ContainerSupport.h:78-90) -> decompose ->
// group -> Kernel.  For a sum of element energies the result always has the same structure:
// one kernel per element shape computing the element's Hessian block, and output kernels that
// sum element entries (plus constants) into every nonzero.  HessianAssembly produces exactly
// that Program directly from the element shapes and their arguments: each element energy is
// analysed (its shape and leaves) and released, the Hessian template is differentiated once
// per shape, and the sparse pattern and contribution lists come from a counting sort of the
// (row, col) pairs.  No per-nonzero expression is ever built, which is what makes the batched
// workload (hundreds of cloths, ~10^7 elements) fit in memory and time.
//
// Several instances (independent variable vectors) can be assembled into one program: instance
// k reads input container k and writes output container k (the full symmetric Hessian in CSC).
#pragma once

#include <cstdint>
#include <vector>

#include "Fused.hpp"
#include "Program.hpp"
#include "Sparse.hpp"
#include "Symbolic.hpp"

namespace Sym {

// Variable groups of element shapes: local variables {k, ELEMENT_LOCAL} (differentiated) and
// parameters {j, ELEMENT_PARAM} (per-element constants).
constexpr int ELEMENT_LOCAL = 249;
constexpr int ELEMENT_PARAM = 248;

class HessianAssembly {
public:
    // Start a new instance whose variables are {0..numVars-1, group}; returns its index.
    // `dofs` > 1 declares the node layout variable = node + dof * (numVars / dofs) used by the
    // fused form (an n x dofs column-major position matrix).
    int addInstance(int numVars, int group, int dofs = 1);
    // Element energy of the current instance.
    void addElement(const Symbolic& energy);

    // Element shape: an energy written over local variables {k, ELEMENT_LOCAL}, k < numLocals,
    // and parameters {j, ELEMENT_PARAM}, j < numParams.  Differentiated once; returns its id.
    int defineShape(const Symbolic& energy, int numLocals, int numParams);
    // Element shape given by its symmetric k x k element matrix instead of an energy (assembly
    // operators such as the cotangent Laplacian): entries[i * k + j] over locals
    // {q, ELEMENT_LOCAL}, q < numLocals, and parameters {j, ELEMENT_PARAM}, j < numParams;
    // entries[i * k + j] and entries[j * k + i] must be the same expression.  Row and column i
    // of the matrix is the variable of local patternLocals[i]; the other locals are inputs the
    // entries read.  Returns its id (used like a shape of defineShape).
    int defineMatrixShape(int k, const Symbolic* entries, const int* patternLocals, int numLocals, int numParams);
    // Element of the current instance with shape `shape`: local k is variable vars[k] of the
    // instance, parameter j has value params[j].  Builds no expression.
    void addElement(int shape, const int* vars, const double* params);
    // The current instance's output is its first dof block only: an n x n CSC (n = numVars /
    // dofs) whose rows and columns are nodes.  Every element row and column (and constant)
    // must then be a dof-0 variable.
    void setNodePattern();
    // Constant Hessian contribution (row, col) of the current instance (not mirrored).
    void addConstant(int row, int col, double value);
    // Constant block on the nodes of the current instance (dofs > 1): for every dof d,
    // H(nodes[i] + d n, nodes[j] + d n) += M[i * numNodes + j], n = numVars / dofs.
    // The isotropic form of quadratic energies such as discrete bending; stored compactly.
    void addNodalConstant(int numNodes, const int* nodes, const double* M);

    size_t numInstances() const { return inst_.size(); }
    size_t numElements() const { return elemTmpl_.size(); }

    // Build the program (inputs: one container per instance; outputs: one CSC per instance).
    Program build(const BuildOptions& opt, ProgramInputs& io, std::vector<SparseMatrix<double>>& patterns) const;
    // Build the fused form (same inputs, outputs and patterns).  Throws std::invalid_argument
    // when an element template has no common node layout or an output has more than 255
    // element contributions.
    FusedProgram buildFused(const BuildOptions& opt, const FusedOptions& fopt, ProgramInputs& io,
                            std::vector<SparseMatrix<double>>& patterns) const;

private:
    struct Instance {
        int numVars = 0, group = 0, dofs = 1;
        bool nodePattern = false;
        size_t firstElem = 0, numElems = 0;
        std::vector<int> cRow, cCol;
        std::vector<double> cVal;
        std::vector<int> ncNodes;          // nodal constant blocks: nodes
        std::vector<size_t> ncStart{0};    // block b: nodes [ncStart[b], ncStart[b+1])
        std::vector<double> ncVal;         // block b: values from ncValStart[b], numNodes^2
        std::vector<size_t> ncValStart{0};
    };
    struct Arg {
        double value;  // constant value
        int32_t id;    // variable id (kind 1)
        int32_t kind;  // 0 constant, 1 variable of this instance
    };
    struct Shape {
        uint32_t tmpl;
        int numLocals, numParams;
        std::vector<int> diffLocal;   // template diff slot -> local variable
        std::vector<int> otherLocal;  // template other slot -> local variable (-1: none)
        std::vector<int> otherParam;  // template other slot -> parameter (-1: constant below)
        std::vector<double> otherConst;
    };
    // Output pattern of one instance in CSC order; contributions are (element << 16 | slot).
    struct InstancePattern {
        std::vector<uint32_t> first, count;  // per output, into contrib
        std::vector<double> cval;            // per output, sum of constants
        std::vector<char> hasConst;
        std::vector<uint64_t> contrib;
        SparseMatrix<double> pattern;
    };
    void instancePattern(size_t k, InstancePattern& out) const;
    // buildFused with block sums (blocks) or one contribution list per output; with blocks it
    // throws when the Hessian is not block-structured
    FusedProgram buildFusedForm(const BuildOptions& opt, const FusedOptions& fopt, ProgramInputs& io,
                                std::vector<SparseMatrix<double>>& patterns, bool blocks) const;
    // the pattern an instance reports (its first dof block for node patterns)
    SparseMatrix<double> reportedPattern(size_t k, SparseMatrix<double>&& full) const;
    void inputLayout(ProgramInputs& io, std::vector<size_t>& inOff) const;

    std::vector<Shape> shapes_;
    std::vector<Instance> inst_;
    std::vector<uint32_t> elemTmpl_;
    std::vector<uint64_t> elemArg_;  // first argument (diff args first, then others)
    std::vector<Arg> args_;
};

}  // namespace Sym
