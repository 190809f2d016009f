// Evaluation and symbolic differentiation (src/SymbolicDifferentiation.hpp:16-290 and
// src/SymbolicDifferentiation.cpp:12-44 of the reference).
//  * evaluate: tree walk, variables {id, group} read values[id];
//  * differentiate: reverse mode, returns d expr / d vars;
//  * hessianCached: forward-over-forward differentiation through a hash-consing factory;
//  * hessianSparse: for a sum of element terms, the upper-triangle (i <= j) Hessian entries of
//    every term, each term's entries sealed into one FIXED block (sealBlock) so that one element
//    becomes one kernel instance.
//    Elements with the same shape share one differentiated template (Template.hpp).
#pragma once

#include <tuple>
#include <unordered_map>
#include <vector>

#include "ContainerSupport.hpp"
#include "Symbolic.hpp"
#include "Traverse.hpp"

namespace Sym {

// Every value of the given containers of Symbolic (anything with contPtr/contSize), in order.
inline void appendSymbols(std::vector<Symbolic>&) {}
template <class C, class... Rest>
void appendSymbols(std::vector<Symbolic>& out, const C& c, const Rest&... rest) {
    const Symbolic* p = contPtr(c);
    out.insert(out.end(), p, p + contSize(c));
    appendSymbols(out, rest...);
}
template <class... Cs>
std::vector<Symbolic> symbolsOf(const Cs&... cs) {
    std::vector<Symbolic> out;
    appendSymbols(out, cs...);
    return out;
}

// Seal a matrix of expressions into one indivisible unit of work: every algebraically
// distinct entry is assigned to a fresh variable in one BLOCK, and each entry becomes
// FIXED(its variable, that block), which decomposition keeps whole.
void sealBlock(std::vector<std::vector<Symbolic>>& entries);

template <class Cont>
double evaluate(const Symbolic& x, const Cont& c) {
    auto cptr = contPtr(c);
    return traverseGenerate<double>(x, [=](const Symbolic& y, std::vector<double>& childVals) {
        if (y.op() == CONST) return y.constant();
        if (y.op() == VAR) return (double)cptr[y.variable()[0]];
        return evaluate(y.op(), childVals);
    });
}

// Evaluate with one value array per variable group (group -> values).
double evaluateGroups(const Symbolic& x, const std::unordered_map<int, const double*>& groups);

std::vector<Symbolic> differentiate(const Symbolic& x, const std::vector<Symbolic>& vars);

template <class... Args>
std::vector<Symbolic> differentiate(const Symbolic& x, Args&... args) {
    return differentiate(x, symbolsOf(args...));
}

std::vector<std::vector<Symbolic>> hessianCached(const Symbolic& x, const std::vector<Symbolic>& vars);
std::vector<Symbolic> gradientCached(const Symbolic& x, const std::vector<Symbolic>& vars);

// Hessian tuples (row, col >= row... by variable index, expression) of a sum of element terms.
// `shareTemplates` = false materialises one expression tree per element, as the reference does.
std::vector<std::tuple<int, int, Symbolic>> hessianSparseVars(const Symbolic& expr, const std::vector<Symbolic>& vars,
                                                              bool shareTemplates = true);

template <class... Args>
std::vector<std::tuple<int, int, Symbolic>> hessianSparse(const Symbolic& expr, Args&... args) {
    return hessianSparseVars(expr, symbolsOf(args...), true);
}

template <class... Args>
std::vector<std::tuple<int, int, Symbolic>> hessianSparseDirect(const Symbolic& expr, Args&... args) {
    return hessianSparseVars(expr, symbolsOf(args...), false);
}

}  // namespace Sym
