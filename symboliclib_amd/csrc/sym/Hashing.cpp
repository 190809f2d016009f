#include "Hashing.hpp"

namespace Sym {

namespace {

// Reflected form of the CRC-64/Jones polynomial.
constexpr uint64_t kPoly = 0x95AC9329AC4BC9B5ULL;

struct CrcTable {
    uint64_t t[256];
    CrcTable() {
        for (unsigned i = 0; i < 256; ++i) {
            uint64_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1) ? (c >> 1) ^ kPoly : (c >> 1);
            t[i] = c;
        }
    }
};

const CrcTable& table() {
    static const CrcTable tab;
    return tab;
}

}  // namespace

uint64_t crc64TableEntry(unsigned i) { return table().t[i & 255]; }

hash_t CRC64_hash(const void* data, size_t len) {
    const uint64_t* t = table().t;
    const unsigned char* s = static_cast<const unsigned char*>(data);
    uint64_t crc = 0;
    for (size_t j = 0; j < len; ++j) crc = t[(uint8_t)crc ^ s[j]] ^ (crc >> 8);
    return (hash_t)crc;
}

hash_t hash(const std::array<int, 2>& x) {
    // src/Hashing.cpp:153-160, with the table index reduced into range for negative ids.
    hash_t data[2];
    for (int k = 0; k < 2; ++k) {
        const int r = ((x[k] % 255) + 255) % 255;
        data[k] = (hash_t)(crc64TableEntry(r + 1) + (uint64_t)(long long)x[k]);
    }
    return CRC64_hash(data, sizeof(data));
}

}  // namespace Sym
