// On-disk form of a built program (the job of the reference's table dump: init() writes
// constData / indexData / outIndexData into the working directory and the generated library
// reads them back, src/ComputeUnit.cpp:364-366, src/StringResources.h:233-269).
//
// One file holds everything a ComputeUnit needs after its front-end and build: the device
// form, the input and output layout, the program tables (index, constants, placements), the
// kernel templates (so a loaded fused unit can still be retuned), the emitted source and its
// launches, the tile pools and the output patterns.  Compiled code objects are not in it: the
// runtime's code-object cache (keyed by the source) supplies them, so loading a program and
// compiling it runs neither the front-end, nor the build, nor the compiler.
//
// Format (little-endian): "SYMLIBPG", u32 version, then sections in a fixed order; vectors are
// a u64 element count followed by the raw elements, strings a u64 byte count and the bytes.
// Symbolic templates are one node table (post-order, children before parents): per node an
// op byte, a child count and child node indices, a double for constants, two ints for
// variables.  Readers check the magic, the version, every count against the bytes left and
// every node index, and throw std::runtime_error on anything else.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "Symbolic.hpp"

namespace Sym {

class BinWriter {
public:
    explicit BinWriter(const std::string& path);
    ~BinWriter();
    BinWriter(const BinWriter&) = delete;
    BinWriter& operator=(const BinWriter&) = delete;

    void raw(const void* p, size_t n);
    template <class T>
    void pod(const T& v) {
        raw(&v, sizeof(T));
    }
    template <class T>
    void vec(const std::vector<T>& v) {
        pod<uint64_t>(v.size());
        if (!v.empty()) raw(v.data(), sizeof(T) * v.size());
    }
    void str(const std::string& s);
    void symbolics(const std::vector<Symbolic>& roots);
    // flush and close; throws when any write failed
    void finish();

private:
    std::FILE* f_ = nullptr;
    std::string path_;
    bool ok_ = true;
};

class BinReader {
public:
    explicit BinReader(const std::string& path);
    ~BinReader();
    BinReader(const BinReader&) = delete;
    BinReader& operator=(const BinReader&) = delete;

    void raw(void* p, size_t n);
    template <class T>
    T pod() {
        T v;
        raw(&v, sizeof(T));
        return v;
    }
    template <class T>
    std::vector<T> vec() {
        const uint64_t n = pod<uint64_t>();
        if (n > left() / sizeof(T)) fail("vector longer than the file");
        std::vector<T> v((size_t)n);
        if (n) raw(v.data(), sizeof(T) * (size_t)n);
        return v;
    }
    std::string str();
    std::vector<Symbolic> symbolics();
    [[noreturn]] void fail(const std::string& what) const;

private:
    uint64_t left() const { return size_ - pos_; }
    std::FILE* f_ = nullptr;
    std::string path_;
    uint64_t size_ = 0, pos_ = 0;
};

constexpr char PROGRAM_FILE_MAGIC[9] = "SYMLIBPG";
constexpr uint32_t PROGRAM_FILE_VERSION = 2;

}  // namespace Sym
