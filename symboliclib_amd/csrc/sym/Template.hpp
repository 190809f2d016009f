// Expression templates: one symbolic block of outputs written over placeholder arguments,
// shared by every element whose expression has the same shape.
//
// The reference builds one full expression tree per element (e.g. one 9x9 Hessian per cloth
// face through src/SymbolicDifferentiation.hpp:234-278) and only later rediscovers, in
// Kernel's lock-step traversal (src/ComputeKernel.cpp:27-268), that all elements share a
// template.  At 10^5-10^7 elements those trees do not fit in host memory, so this build
// differentiates one representative per shape and represents every element by an INSTANCE
// node: children = [output variables..., arguments...], tag = template id.  Kernel expands
// the instance back into the template's block once per group, so the emitted code is the
// same as for fully materialised trees.
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <unordered_map>
#include <vector>

#include "Symbolic.hpp"

namespace Sym {

struct ExprTemplate {
    uint32_t id = 0;
    int numArgs = 0;       // argument slots, referenced as VAR(slot, TEMPLATE_ARG)
    int numOutputs = 0;    // output expressions
    std::vector<Symbolic> outputs;
    std::vector<int64_t> signature;  // shape key the template was built for
    hash_t signatureHash = 0;
    // Hessian templates: output index of entry (i, j) over the differentiated arguments
    int numDiffVars = 0;
    std::vector<int> entryToOutput;
};

// Registry of all templates of the process (not thread-safe, like the rest of the IR).
uint32_t registerTemplate(std::unique_ptr<ExprTemplate> t);
const ExprTemplate& getTemplate(uint32_t id);
// Template with exactly this signature, or nullptr.
const ExprTemplate* findTemplate(hash_t signatureHash, const std::vector<int64_t>& signature);

// INSTANCE(out_0..out_{m-1}, arg_0..arg_{a-1}) -> BLOCK(ASSIGN(out_k, output_k[args]))
Symbolic expandInstance(const Symbolic& inst);

// Replace VAR(slot, TEMPLATE_ARG) leaves of `x` by args[slot].
Symbolic substituteArguments(const Symbolic& x, const std::vector<Symbolic>& args);

// Shape of one element term: a token string that fixes the DAG (operations, sharing, integer
// literals) up to the identity of its leaves, plus the leaves in slot order.  Variables found
// in `varId` are differentiated arguments (slots 0..nd-1, first-occurrence order); other
// variables and non-integer constants are further arguments (slots nd..).
struct ElementShape {
    std::vector<int64_t> sig;
    std::vector<Symbolic> diffLeaves;
    std::vector<Symbolic> otherLeaves;
    std::vector<Symbolic> nodes;  // every DAG node, post-order
    std::vector<std::vector<int>> childIdx;
    std::vector<int> leafKind;  // 0 interior, 1 differentiated leaf, 2 other argument, 3 literal
    std::vector<int> leafSlot;
};

bool isLiteralConstant(double c);
void analyzeShape(const Symbolic& root, const std::unordered_map<hash_t, size_t>& varId, ElementShape& s);
void analyzeShape(const Symbolic& root, const std::function<bool(const Symbolic&)>& isDiff, ElementShape& s);
// Hessian template of the element shape over its differentiated arguments (built once per shape).
const ExprTemplate& hessianTemplateFor(const ElementShape& s);

}  // namespace Sym
