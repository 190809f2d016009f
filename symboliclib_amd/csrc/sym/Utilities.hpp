// Small container and graph helpers used by the compiler middle-end
// (src/Utilities.hpp:54-200 of the reference).
#pragma once

#include <algorithm>
#include <cassert>
#include <cstddef>
#include <fstream>
#include <iostream>
#include <iterator>
#include <limits>
#include <string>
#include <vector>

namespace Sym {

// Kahn's algorithm over an adjacency list graph[node] = successors.  Returns an empty order
// (and reports) when the graph has a cycle.
template <class IdCont>
std::vector<int> topologicalOrder(const std::vector<IdCont>& graph) {
    const size_t n = graph.size();
    std::vector<int> incoming(n, 0);
    for (auto& nbh : graph)
        for (int i : nbh) ++incoming[i];

    std::vector<int> S, L;
    L.reserve(n);
    for (size_t i = 0; i < n; ++i)
        if (!incoming[i]) S.push_back((int)i);

    while (!S.empty()) {
        const int node = S.back();
        S.pop_back();
        L.push_back(node);
        for (auto m : graph[node])
            if (--incoming[m] == 0) S.push_back(m);
    }

    for (auto i : incoming)
        if (i) {
            std::cerr << "topological ordering failed: dependency cycle" << std::endl;
            return {};
        }
    return L;
}

template <class Cont, class IdCont>
void reorder(Cont& c, const IdCont& ids) {
    Cont ret;
    ret.reserve(ids.size());
    for (size_t i = 0; i < ids.size(); ++i) ret.push_back(std::move(c[ids[i]]));
    c.swap(ret);
}

// Append the columns of `src` (all of equal length) to `dst`.
//   packLength 0: column after column (structure of arrays)
//   packLength 1: row after row
//   packLength k: blocks of k rows, each block column after column
// Returns the pack length used (0 when the length is not a multiple of k).
template <class T>
int appendTransposed(const std::vector<std::vector<T>>& src, std::vector<T>& dst, const size_t packLength = 1) {
    if (src.empty()) return (int)packLength;
    const size_t segLength = src.front().size();
    for (auto& x : src) {
        (void)x;
        assert(x.size() == segLength);
    }
    const size_t n0 = dst.size();
    dst.resize(n0 + segLength * src.size());
    auto ptr = dst.begin() + n0;
    if (packLength == 0 || segLength % packLength) {
        for (const auto& s : src) ptr = std::copy(s.begin(), s.end(), ptr);
        return 0;
    }
    const size_t packs = segLength / packLength;
    for (size_t i = 0; i < packs; ++i)
        for (auto& s : src) ptr = std::copy_n(s.begin() + packLength * i, packLength, ptr);
    return (int)packLength;
}

template <class RandomAccessIterator, class Key>
void sortBy(RandomAccessIterator first, RandomAccessIterator last, Key key) {
    const auto d = std::distance(first, last);
    if (d < 2) return;
    if (d == 2) {
        if (key(*first) > key(*(first + 1))) std::iter_swap(first, first + 1);
    } else {
        std::sort(first, last, [&](const auto& x, const auto& y) { return key(x) < key(y); });
    }
}

template <class TCont>
void deleteDuplicates(TCont& cont) {
    std::sort(cont.begin(), cont.end());
    cont.erase(std::unique(cont.begin(), cont.end()), cont.end());
}

template <class T, class TCont2>
void append(std::vector<T>& cont, const TCont2& cont2) {
    cont.insert(cont.end(), std::begin(cont2), std::end(cont2));
}

template <class Cont, class Func>
auto mapData(const Cont& c, Func f) {
    std::vector<std::decay_t<decltype(f(*std::begin(c)))>> ret;
    ret.reserve((size_t)std::distance(std::begin(c), std::end(c)));
    for (const auto& x : c) ret.push_back(f(x));
    return ret;
}

template <class T>
std::vector<T> flatten(const std::vector<std::vector<T>>& cont) {
    size_t sz = 0;
    for (auto& c : cont) sz += c.size();
    std::vector<T> ret;
    ret.reserve(sz);
    for (auto& c : cont) ret.insert(ret.end(), c.begin(), c.end());
    return ret;
}

template <class T>
void saveData(const std::string& fname, const std::vector<T>& c) {
    std::ofstream file(fname, std::ios::out | std::ios::binary);
    file.write((const char*)c.data(), c.size() * sizeof(T));
}

}  // namespace Sym
