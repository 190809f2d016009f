// Hashing used by the symbolic IR: a 64-bit CRC over raw bytes plus typed helpers.
//
// Follows src/Hashing.hpp:8-34 and src/Hashing.cpp:142-170 of the reference: the same
// reflected CRC-64 (Jones polynomial), the same "hash of the bytes of a value" helpers and
// the same small-constant test (|h| >> 4 == 0).  The table is generated from the polynomial
// at start-up instead of being spelled out.
//
// One deliberate difference: the reference's variable hash indexes its table with
// `x % 255 + 1`, which reads outside the table for negative variable groups (the template
// tables use groups -2, -3, -4).  Here the index is reduced into [1, 255] first, so every
// variable hash is well defined.  Hash values only decide sub-expression identity and the
// order of commutative operands, never an output value or pattern.
#pragma once

#include <array>
#include <cstdint>
#include <cstring>
#include <vector>

namespace Sym {

typedef long long hash_t;

// Table-driven reflected CRC-64 over `len` bytes.
hash_t CRC64_hash(const void* data, size_t len);

// i-th entry (0..255) of the CRC table; exposed for the variable hash.
uint64_t crc64TableEntry(unsigned i);

template <class T>
hash_t hash(const std::vector<T>& data) {
    return CRC64_hash(data.data(), sizeof(T) * data.size());
}

template <class T1, class T2>
hash_t hash(const T1& x1, const T2& x2) {
    unsigned char buf[sizeof(T1) + sizeof(T2)];
    std::memcpy(buf, &x1, sizeof(T1));
    std::memcpy(buf + sizeof(T1), &x2, sizeof(T2));
    return CRC64_hash(buf, sizeof(buf));
}

// hash of a variable {id, group}
hash_t hash(const std::array<int, 2>& x);

template <class T>
hash_t hash(const T& data) {
    return CRC64_hash(&data, sizeof(T));
}

// Expressions whose algebraic hash is in (-16, 16) are treated as that integer constant.
inline bool isSmallConstant(hash_t h) {
    if (h < 0) h = -h;
    return (h >> 4) == 0;
}

}  // namespace Sym
