// Open-addressing hash map keyed by non-zero 64-bit node identities (pointer values).
// Used where the compiler passes index millions of nodes and std::unordered_map's per-node
// allocations dominate.
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

namespace Sym {

template <class V>
class IdMap {
    std::vector<long long> keys_;
    std::vector<V> vals_;
    size_t mask_ = 0, count_ = 0;

    static size_t mix(long long k) {
        uint64_t x = (uint64_t)k;
        x ^= x >> 33;
        x *= 0xff51afd7ed558ccdULL;
        x ^= x >> 33;
        return (size_t)x;
    }

    void grow() {
        std::vector<long long> ok;
        std::vector<V> ov;
        ok.swap(keys_);
        ov.swap(vals_);
        const size_t cap = ok.empty() ? 64 : ok.size() * 2;
        keys_.assign(cap, 0);
        vals_.assign(cap, V{});
        mask_ = cap - 1;
        count_ = 0;
        for (size_t i = 0; i < ok.size(); ++i)
            if (ok[i]) (*this)[ok[i]] = std::move(ov[i]);
    }

public:
    explicit IdMap(size_t expected = 0) {
        size_t cap = 64;
        while (cap < 2 * expected) cap <<= 1;
        keys_.assign(cap, 0);
        vals_.assign(cap, V{});
        mask_ = cap - 1;
    }

    size_t size() const { return count_; }

    V& operator[](long long k) {
        if (2 * (count_ + 1) > keys_.size()) grow();
        size_t i = mix(k) & mask_;
        while (keys_[i] && keys_[i] != k) i = (i + 1) & mask_;
        if (!keys_[i]) {
            keys_[i] = k;
            ++count_;
        }
        return vals_[i];
    }

    V* find(long long k) {
        size_t i = mix(k) & mask_;
        while (keys_[i]) {
            if (keys_[i] == k) return &vals_[i];
            i = (i + 1) & mask_;
        }
        return nullptr;
    }
};

}  // namespace Sym
