#include "ComputeKernel.hpp"

#include <map>
#include <stdexcept>
#include <unordered_set>

#include "Decomposition.hpp"
#include "Simplify.hpp"
#include "Template.hpp"
#include "Traverse.hpp"
#include "Utilities.hpp"

namespace Sym {

namespace {

hash_t hashSymbolic(const std::vector<Symbolic>& x) {
    std::vector<hash_t> hashes;
    hashes.reserve(x.size());
    for (auto& xi : x) hashes.push_back(xi.ahash());
    return hash(hashes);
}

// The same child position of every instance, visited together.
struct StackData {
    std::vector<Symbolic> x;
    hash_t h = 0;
    unsigned childCounter = 0;

    StackData(std::vector<Symbolic>&& data) : x(std::move(data)) { h = hashSymbolic(x); }

    StackData(const std::vector<Symbolic>& parents, unsigned k) {
        x.reserve(parents.size());
        for (auto& d : parents) x.push_back(d[k]);
        h = hashSymbolic(x);
    }

    bool visit() { return childCounter++ == numChilds(); }
    unsigned numChilds() const { return x.empty() ? 0 : x.front().numChilds(); }
    OpType op() const { return x.empty() ? NOOP : x.front().op(); }

    bool isValid() const {
        const OpType o = op();
        const unsigned nc = numChilds();
        const uint32_t tag = x.empty() ? 0 : x.front().tag();
        for (auto& y : x)
            if (y.op() != o || y.numChilds() != nc || (o == INSTANCE && y.tag() != tag)) return false;
        return true;
    }

    // per-instance variables of this position (never collapsed: every instance keeps its entry)
    std::vector<std::array<int, 2>> getVariables() const {
        std::vector<std::array<int, 2>> vals(x.size());
        for (size_t i = 0; i < x.size(); ++i) vals[i] = x[i].variable();
        return vals;
    }

    // per-instance constants; a single value when all instances agree
    std::vector<double> getConstants() const {
        std::vector<double> vals(x.size());
        bool allIdentical = true;
        for (size_t i = 0; i < x.size(); ++i) {
            vals[i] = x[i].constant();
            if (allIdentical && !(vals[0] == vals[i])) allIdentical = false;
        }
        if (allIdentical) vals.resize(1);
        return vals;
    }
};

}  // namespace

Kernel::Kernel(const std::vector<Symbolic>& expressions, const int simdWidth_, const bool simplifyTemplate)
    : n(expressions.size()), simdWidth(simdWidth_) {
    if (expressions.empty()) throw std::invalid_argument("Kernel: empty expression group");

    std::map<hash_t, Symbolic> xprMap;
    std::vector<StackData> stack;
    stack.emplace_back(std::vector<Symbolic>(expressions));
    std::vector<Symbolic> tStack;

    while (!stack.empty()) {
        StackData& curr = stack.back();
        if (!curr.isValid()) throw std::logic_error("Kernel: instances of a group differ in structure");

        if (curr.visit()) {
            const unsigned nc = curr.numChilds();
            std::vector<Symbolic> ch(tStack.end() - nc, tStack.end());
            tStack.erase(tStack.end() - nc, tStack.end());
            Symbolic txpr = curr.op() == INSTANCE ? Symbolic::instance(curr.x.front().tag(), ch, 0) : Symbolic(curr.op(), ch);
            tStack.push_back(txpr);
            if (xprMap.find(curr.h) == xprMap.end()) xprMap[curr.h] = txpr;
            stack.pop_back();
        } else {
            const unsigned k = curr.childCounter - 1;
            StackData data(curr.x, k);
            auto it = xprMap.find(data.h);
            if (it != xprMap.end()) {
                tStack.push_back(it->second);
            } else if (data.op() == VAR) {
                bool isOutput = curr.op() == ASSIGN && k == 0;
                if (curr.op() == INSTANCE) isOutput = (int)k < getTemplate(curr.x.front().tag()).numOutputs;
                if (isOutput) {
                    tStack.push_back(Symbolic((int)outputVariableTable.size(), -2));
                    outputVariableTable.push_back(data.getVariables());
                } else {
                    tStack.push_back(Symbolic((int)variableTable.size(), -3));
                    variableTable.push_back(data.getVariables());
                }
                xprMap[data.h] = tStack.back();
            } else if (data.op() == CONST) {
                auto values = data.getConstants();
                if (values.size() == 1) {
                    tStack.push_back(Symbolic(values.front()));
                } else {
                    tStack.push_back(Symbolic((int)constantTable.size(), -4));
                    constantTable.push_back(std::move(values));
                }
                xprMap[data.h] = tStack.back();
            } else {
                stack.push_back(std::move(data));  // invalidates `curr`
            }
        }
    }

    // pad to a multiple of the SIMD width (CPU code only)
    if (simdWidth > 1 && n > (size_t)simdWidth && n % simdWidth) {
        const size_t pad = simdWidth - n % simdWidth;
        for (auto& cnst : constantTable) cnst.insert(cnst.end(), pad, 0.0);
        for (auto& vars : variableTable) vars.insert(vars.end(), pad, vars.back());
        for (auto& vars : outputVariableTable) vars.insert(vars.end(), pad, std::array<int, 2>{-1, -1});
        n += pad;
    }

    assert(tStack.size() == 1);
    templateExpression = tStack.front();
    optimizeTemplateExpression(simplifyTemplate);
}

void Kernel::optimizeTemplateExpression(bool simplifyTemplate) {
    templateExpression = optimizeTemplate(templateExpression, simplifyTemplate);
}

Symbolic optimizeTemplate(Symbolic templateExpression, bool simplifyTemplate) {
    if (templateExpression.op() == INSTANCE) templateExpression = expandInstance(templateExpression);

    // fold constants; algebraically equal sub-expressions become one shared node
    templateExpression = removeConstantExpressions(templateExpression);

    std::vector<Symbolic> vec{templateExpression};
    auto blocks = decompose(vec, 0, LOCAL_VAR, false);

    if (simplifyTemplate)
        for (auto& b : blocks) b.x = simplify(b.x);

    for (auto& b : blocks)
        if (b.x.op() == ASSIGN && b.x[0].op() == VAR && b.x[0].variable()[1] == LOCAL_VAR)
            b.x = Symbolic(DEFINE, b.x[0], b.x[1]);

    // definitions before their uses
    std::vector<std::unordered_set<int>> graph(blocks.size());
    for (size_t i = 0; i < blocks.size(); ++i)
        for (auto& c : blocks[i].childs) graph[c].insert((int)i);

    auto order = topologicalOrder(graph);
    if (order.size() != blocks.size()) throw std::logic_error("Kernel: cyclic local definitions");
    std::vector<Symbolic> kernelExpressions;
    kernelExpressions.reserve(order.size());
    for (int i : order) kernelExpressions.push_back(blocks[i].x);
    return Symbolic(BLOCK, kernelExpressions);
}

int Kernel::getPackLength() {
    getConstantData();
    getOutputVariableData();
    getVariableData();
    assert(constantsPackLength == variablePackLength && constantsPackLength == outputVariablePackLength);
    return variablePackLength;
}

const std::vector<double>& Kernel::getConstantData() {
    if (constantsPackLength < 0) constantsPackLength = appendTransposed(constantTable, constantData, simdWidth);
    return constantData;
}

const std::vector<std::array<int, 2>>& Kernel::getOutputVariableData() {
    if (outputVariablePackLength < 0)
        outputVariablePackLength = appendTransposed(outputVariableTable, outputVariableData, simdWidth);
    return outputVariableData;
}

const std::vector<std::array<int, 2>>& Kernel::getVariableData() {
    if (variablePackLength < 0) {
        findCorrelatedVariables();
        variablePackLength = appendTransposed(variableTable, variableData, simdWidth);
    }
    return variableData;
}

const std::vector<Kernel::VariableAccessData>& Kernel::getVariableAccess() {
    findCorrelatedVariables();
    return variableAccessData;
}

size_t Kernel::numVariableSlots() {
    findCorrelatedVariables();
    return variableTable.size();
}

// A column whose entries equal another column's plus one constant (same input group in
// every instance) is not stored; its leaves read the other column with that offset.
void Kernel::findCorrelatedVariables() {
    if (correlated) return;
    correlated = true;

    std::map<hash_t, std::vector<int>> varMap;  // pattern hash -> candidate base columns
    std::vector<std::vector<int>> patterns(variableTable.size());
    for (size_t i = 0; i < variableTable.size(); ++i) {
        const auto& vars = variableTable[i];
        bool flag = true;
        std::vector<int>& vals = patterns[i];
        vals.push_back(vars[0][1]);
        for (auto& v : vars) {
            if (v[1] != vars[0][1]) {
                flag = false;
                break;
            }
            vals.push_back(v[0] - vars[0][0]);
        }
        if (!flag || vars[0][1] == GLOBAL_INTERMEDIATE || vars[0][1] == EXPLICIT_INTERMEDIATE) {
            variableAccessData.push_back({(int)i, 0});
            continue;
        }
        const hash_t h = hash(vals);
        auto& cands = varMap[h];
        int found = -1;
        for (int c : cands)
            if (patterns[c] == vals) {
                found = c;
                break;
            }
        if (found < 0) {
            cands.push_back((int)i);
            variableAccessData.push_back({(int)i, 0});
        } else {
            variableAccessData.push_back({found, vars[0][0] - variableTable[found][0][0]});
        }
    }

    // keep only base columns
    std::map<int, int> varIds;
    std::vector<std::vector<std::array<int, 2>>> variableTable2;
    for (auto& va : variableAccessData) {
        auto it = varIds.find(va.baseVariableId);
        if (it == varIds.end()) {
            const int nid = (int)variableTable2.size();
            varIds[va.baseVariableId] = nid;
            variableTable2.push_back(variableTable[va.baseVariableId]);
            va.baseVariableId = nid;
        } else {
            va.baseVariableId = it->second;
        }
    }
    variableTable.swap(variableTable2);
}

}  // namespace Sym
