// A kernel = one template expression evaluated for many instances.
//
// Follows src/ComputeKernel.hpp and src/ComputeKernel.cpp:27-484 of the reference: all
// instances of a structure group are traversed in lock-step; wherever the instances differ,
// the template reads a per-instance table instead:
//   VAR(k, -2)  output variable table k   (the left side of an assignment)
//   VAR(k, -3)  input variable table k    (gathered values)
//   VAR(k, -4)  constant table k          (constants that differ between instances)
// Constants equal in every instance stay literals.  The template then goes through local
// common-sub-expression extraction (LOCAL_VAR definitions) and algebraic simplification.
// Tables are laid out column after column (packLength 0) for the GPU, or in blocks of
// `simdWidth` rows for vectorised CPU code.
#pragma once

#include <array>
#include <vector>

#include "Symbolic.hpp"

namespace Sym {

class Kernel {
public:
    struct VariableAccessData {
        int baseVariableId;  // table column actually stored
        int offset;          // constant added to the stored index
    };

private:
    size_t n;
    int simdWidth;
    Symbolic templateExpression;
    size_t globalOutputVariableOffset = 0;
    size_t globalVariableOffset = 0;
    size_t globalConstantsOffset = 0;

    int outputVariablePackLength = -1;
    std::vector<std::array<int, 2>> outputVariableData;
    std::vector<std::vector<std::array<int, 2>>> outputVariableTable;

    int constantsPackLength = -1;
    std::vector<double> constantData;
    std::vector<std::vector<double>> constantTable;

    int variablePackLength = -1;
    std::vector<std::array<int, 2>> variableData;
    std::vector<std::vector<std::array<int, 2>>> variableTable;
    std::vector<VariableAccessData> variableAccessData;
    bool correlated = false;

    void findCorrelatedVariables();

public:
    Kernel(const std::vector<Symbolic>& expressions, const int simdWidth = 1, const bool simplifyTemplate = true);

    void optimizeTemplateExpression(bool simplifyTemplate);

    int getPackLength();
    void setGlobalOutputOffset(const size_t offset) { globalOutputVariableOffset = offset; }
    void setGlobalIndexOffset(const size_t offset) { globalVariableOffset = offset; }
    void setGlobalConstantsOffset(const size_t offset) { globalConstantsOffset = offset; }
    size_t getGlobalOutputOffset() const { return globalOutputVariableOffset; }
    size_t getGlobalIndexOffset() const { return globalVariableOffset; }
    size_t getGlobalConstantsOffset() const { return globalConstantsOffset; }

    int numInstances() const { return (int)n; }
    size_t numOutputSlots() const { return outputVariableTable.size(); }
    size_t numConstantSlots() const { return constantTable.size(); }
    size_t numVariableSlots();  // after correlation compression

    const std::vector<double>& getConstantData();
    const std::vector<std::array<int, 2>>& getOutputVariableData();
    const std::vector<std::array<int, 2>>& getVariableData();
    const std::vector<VariableAccessData>& getVariableAccess();
    const std::vector<std::vector<std::array<int, 2>>>& getOutputVariableTable() const { return outputVariableTable; }

    Symbolic getTemplateExpression() const { return templateExpression; }
};

// Template optimisation shared by Kernel and the direct assembly builder
// (src/ComputeKernel.cpp:270-323): expand an INSTANCE, fold constants and merge algebraically
// equal sub-expressions, extract local definitions (LOCAL_VAR), simplify each block, order
// definitions before uses.  Returns a BLOCK.
Symbolic optimizeTemplate(Symbolic templateExpression, bool simplifyTemplate);

}  // namespace Sym
