// Helpers over containers of symbolic expressions (src/SymbolicUtilities.hpp:10-86 of the
// reference): naming variables, flattening several containers into one list, and makeFixed,
// which turns a set of expressions into one indivisible block of assignments.
#pragma once

#include <unordered_map>
#include <utility>
#include <vector>

#include "ContainerSupport.hpp"
#include "Symbolic.hpp"
#include "Traverse.hpp"

namespace Sym {

template <class Cont>
void setVariables(Cont& A, const int objId) {
    for (size_t i = 0; i < contSize(A); ++i) contPtr(A)[i] = Symbolic((int)i, objId);
}

template <class Cont>
void assignToVariables(Cont& A, const int objId) {
    for (size_t i = 0; i < contSize(A); ++i) contPtr(A)[i] = Symbolic(ASSIGN, Symbolic((int)i, objId), contPtr(A)[i]);
}

inline void addSymbolic(Symbolic* x, size_t s, std::vector<std::pair<Symbolic*, size_t>>& data) {
    data.push_back(std::make_pair(x, s));
}
inline void addSymbolic(const Symbolic* x, size_t s, std::vector<std::pair<Symbolic*, size_t>>& data) {
    data.push_back(std::make_pair(const_cast<Symbolic*>(x), s));
}
template <class T>
void addSymbolic(T*, size_t, std::vector<std::pair<Symbolic*, size_t>>&) {}

inline void flattenContainerData(std::vector<std::pair<Symbolic*, size_t>>&) {}

template <class... TArgs>
void flattenContainerData(std::vector<std::pair<Symbolic*, size_t>>& data, std::vector<std::vector<Symbolic>>& arg,
                          TArgs&... args);

template <class TArg, class... TArgs>
void flattenContainerData(std::vector<std::pair<Symbolic*, size_t>>& data, TArg& arg, TArgs&... args) {
    addSymbolic(contPtr(arg), contSize(arg), data);
    flattenContainerData(data, args...);
}

template <class... TArgs>
void flattenContainerData(std::vector<std::pair<Symbolic*, size_t>>& data, std::vector<std::vector<Symbolic>>& arg,
                          TArgs&... args) {
    for (auto& a : arg) addSymbolic(contPtr(a), contSize(a), data);
    flattenContainerData(data, args...);
}

template <class... TArgs>
std::vector<Symbolic> flattenContainerData(TArgs&... args) {
    std::vector<std::pair<Symbolic*, size_t>> data;
    flattenContainerData(data, args...);
    std::vector<Symbolic> ret;
    for (auto& d : data) ret.insert(ret.end(), d.first, d.first + d.second);
    return ret;
}

// Replace every expression e by FIXED(v_e, B) where B = BLOCK(ASSIGN(v_e, e), ...) holds one
// assignment per algebraically distinct expression and v_e is a fresh variable.
template <class... TArgs>
void makeFixed(TArgs&... args) {
    std::vector<std::pair<Symbolic*, size_t>> data;
    flattenContainerData(data, args...);
    if (data.empty()) return;

    std::unordered_map<hash_t, Symbolic, IdentityHash<hash_t>> exprMap;
    std::vector<Symbolic> expressions;
    for (auto& d : data)
        for (size_t i = 0; i < d.second; ++i) {
            auto& v = exprMap[d.first[i].ahash()];
            if (v.op() == NOOP) {
                v = Symbolic::generateUniqueVariable();
                expressions.emplace_back(ASSIGN, v, d.first[i]);
            }
        }

    Symbolic block(BLOCK, expressions);
    for (auto& d : data)
        for (size_t i = 0; i < d.second; ++i) d.first[i] = Symbolic(FIXED, exprMap[d.first[i].ahash()], block);
}

}  // namespace Sym
