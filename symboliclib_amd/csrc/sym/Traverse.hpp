// Iterative traversals of expression DAGs (src/Traverse.hpp:18-74, src/Traverse.cpp:11-77 of
// the reference).  Unlike the reference, every traversal owns its stack, so traversals may
// nest and the functions are reentrant.
#pragma once

#include <functional>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

#include "Symbolic.hpp"

namespace Sym {

template <class T>
struct IdentityHash {
    inline size_t operator()(const T& h) const { return (size_t)h; }
};

// Pre-order walk; `fun` returns whether to descend.  With `unique`, a node reached twice
// through sharing is visited once.
template <class F>
void preOrderTraverse(const Symbolic& x, F&& fun, const bool unique = false) {
    std::unordered_set<long long> visited;
    thread_local std::vector<const Symbolic*> stackStore;
    std::vector<const Symbolic*> stack;
    stack.swap(stackStore);  // reuse capacity; nested calls get a fresh vector
    stack.clear();
    stack.push_back(&x);
    struct Restore {
        std::vector<const Symbolic*>& s;
        std::vector<const Symbolic*>& store;
        ~Restore() {
            if (s.capacity() > store.capacity()) s.swap(store);
        }
    } restore{stack, stackStore};
    while (!stack.empty()) {
        const Symbolic* y = stack.back();
        stack.pop_back();
        if (unique && !visited.insert(y->id()).second) continue;
        if (fun(*y)) {
            const unsigned nc = y->numChilds();
            if (nc) {
                const Symbolic* ch = y->begin();
                for (unsigned i = 0; i < nc; ++i) stack.push_back(ch + i);
            }
        }
    }
}

// `preFun` decides whether a node's children are visited (and whether postFun runs for it);
// `postFun` runs once all visited children are done.
template <class Pre, class Post>
void prePostOrderTraverse(const Symbolic& x, Pre&& preFun, Post&& postFun) {
    if (!preFun(x)) return;
    std::vector<std::pair<const Symbolic*, unsigned>> stack{{&x, 0u}};
    while (!stack.empty()) {
        const Symbolic* y = stack.back().first;
        const unsigned cnt = stack.back().second;
        const unsigned nchilds = y->numChilds();
        if (nchilds == cnt) {
            postFun(*y);
            stack.pop_back();
        } else {
            ++stack.back().second;
            const Symbolic* c = y->begin() + cnt;
            if (preFun(*c)) stack.push_back({c, 0u});
        }
    }
}

template <class Post>
void postOrderTraverse(const Symbolic& x, Post&& fun) {
    prePostOrderTraverse(x, [](const Symbolic&) { return true; }, fun);
}

std::vector<Symbolic> harvestOp(OpType op, const Symbolic& x);

// Bottom-up generation with memoisation by node identity (or by algebraic hash when
// `hashId`): `postFun(node, childValues)` builds a node's value from its children's values.
template <class T, class Pre, class Post>
T traverseGenerate(const Symbolic& x, Pre&& preFun, Post&& postFun,
                   std::unordered_map<long long, T, IdentityHash<long long>>& valueMap, const bool hashId = false) {
    std::vector<std::pair<Symbolic, unsigned>> stack;
    stack.emplace_back(preFun(x), 0u);
    std::vector<T> valueStack;

    while (!stack.empty()) {
        Symbolic y = stack.back().first;
        const unsigned cnt = stack.back().second;
        const unsigned nchilds = y.numChilds();
        if (nchilds == cnt) {
            std::vector<T> childVals(std::make_move_iterator(valueStack.end() - nchilds),
                                     std::make_move_iterator(valueStack.end()));
            valueStack.erase(valueStack.end() - nchilds, valueStack.end());
            T val = postFun(y, childVals);
            valueMap[hashId ? y.ahash() : y.id()] = val;
            valueStack.push_back(std::move(val));
            stack.pop_back();
        } else {
            ++stack.back().second;
            const Symbolic& c = y[cnt];
            auto it = valueMap.find(hashId ? c.ahash() : c.id());
            if (it != valueMap.end()) valueStack.push_back(it->second);
            else stack.emplace_back(preFun(c), 0u);
        }
    }
    assert(valueStack.size() == 1);
    return std::move(valueStack.front());
}

template <class T, class Pre, class Post>
T traverseGenerate(const Symbolic& x, Pre&& preFun, Post&& postFun, const bool hashId) {
    std::unordered_map<long long, T, IdentityHash<long long>> valueMap;
    return traverseGenerate<T>(x, preFun, postFun, valueMap, hashId);
}

template <class T, class Post>
T traverseGenerate(const Symbolic& x, Post&& postFun) {
    std::unordered_map<long long, T, IdentityHash<long long>> valueMap;
    return traverseGenerate<T>(x, [](const Symbolic& y) { return y; }, postFun, valueMap, false);
}

template <class T, class Post>
T traverseGenerateMap(const Symbolic& x, Post&& postFun,
                      std::unordered_map<long long, T, IdentityHash<long long>>& valueMap, const bool hashId) {
    return traverseGenerate<T>(x, [](const Symbolic& y) { return y; }, postFun, valueMap, hashId);
}

}  // namespace Sym
