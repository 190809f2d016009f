#include "Persist.hpp"

#include <stdexcept>
#include <unordered_map>

namespace Sym {

BinWriter::BinWriter(const std::string& path) : path_(path) {
    f_ = std::fopen(path.c_str(), "wb");
    if (!f_) throw std::runtime_error("cannot write " + path);
}

BinWriter::~BinWriter() {
    if (f_) std::fclose(f_);
}

void BinWriter::raw(const void* p, size_t n) {
    if (n && std::fwrite(p, 1, n, f_) != n) ok_ = false;
}

void BinWriter::str(const std::string& s) {
    pod<uint64_t>(s.size());
    raw(s.data(), s.size());
}

void BinWriter::symbolics(const std::vector<Symbolic>& roots) {
    // post-order node table; a node reached twice is written once
    std::unordered_map<long long, uint32_t> index;
    std::vector<Symbolic> order;
    struct Frame {
        Symbolic y;
        unsigned next;
    };
    for (const Symbolic& r : roots) {
        if (index.count(r.id())) continue;
        std::vector<Frame> st{{r, 0}};
        while (!st.empty()) {
            Frame& f = st.back();
            if (f.next < f.y.numChilds()) {
                const Symbolic c = f.y[f.next++];
                if (!index.count(c.id())) st.push_back({c, 0});
                continue;
            }
            if (!index.count(f.y.id())) {
                if (f.y.op() == INSTANCE) throw std::invalid_argument("program file: unexpanded template instance");
                index.emplace(f.y.id(), (uint32_t)order.size());
                order.push_back(f.y);
            }
            st.pop_back();
        }
    }
    pod<uint64_t>(order.size());
    for (const Symbolic& y : order) {
        pod<uint8_t>((uint8_t)y.op());
        pod<uint32_t>(y.numChilds());
        for (unsigned k = 0; k < y.numChilds(); ++k) pod<uint32_t>(index.at(y[k].id()));
        if (y.op() == CONST) pod<double>(y.constant());
        if (y.op() == VAR) {
            pod<int32_t>(y.variable()[0]);
            pod<int32_t>(y.variable()[1]);
        }
    }
    pod<uint64_t>(roots.size());
    for (const Symbolic& r : roots) pod<uint32_t>(index.at(r.id()));
}

void BinWriter::finish() {
    if (!f_) return;
    const bool closed = std::fclose(f_) == 0;
    f_ = nullptr;
    if (!ok_ || !closed) throw std::runtime_error("writing " + path_ + " failed");
}

BinReader::BinReader(const std::string& path) : path_(path) {
    f_ = std::fopen(path.c_str(), "rb");
    if (!f_) throw std::runtime_error("cannot read " + path);
    if (std::fseek(f_, 0, SEEK_END) != 0) fail("cannot seek");
    const long n = std::ftell(f_);
    if (n < 0 || std::fseek(f_, 0, SEEK_SET) != 0) fail("cannot seek");
    size_ = (uint64_t)n;
}

BinReader::~BinReader() {
    if (f_) std::fclose(f_);
}

void BinReader::fail(const std::string& what) const { throw std::runtime_error("program file " + path_ + ": " + what); }

void BinReader::raw(void* p, size_t n) {
    if (n > left()) fail("truncated");
    if (n && std::fread(p, 1, n, f_) != n) fail("read error");
    pos_ += n;
}

std::string BinReader::str() {
    const uint64_t n = pod<uint64_t>();
    if (n > left()) fail("string longer than the file");
    std::string s((size_t)n, '\0');
    raw(s.data(), (size_t)n);
    return s;
}

std::vector<Symbolic> BinReader::symbolics() {
    const uint64_t n = pod<uint64_t>();
    if (n > left() / 5) fail("node table longer than the file");
    std::vector<Symbolic> nodes;
    nodes.reserve((size_t)n);
    std::vector<Symbolic> childs;
    for (uint64_t i = 0; i < n; ++i) {
        const uint8_t op = pod<uint8_t>();
        const uint32_t nc = pod<uint32_t>();
        if (op >= (uint8_t)NUM_OPS || op == (uint8_t)INSTANCE) fail("unknown operation");
        if (nc > left() / 4) fail("child list longer than the file");
        childs.clear();
        for (uint32_t k = 0; k < nc; ++k) {
            const uint32_t c = pod<uint32_t>();
            if (c >= nodes.size()) fail("child index out of order");
            childs.push_back(nodes[c]);
        }
        if (op == CONST) {
            if (nc) fail("constant with children");
            nodes.emplace_back(pod<double>());
        } else if (op == VAR) {
            if (nc) fail("variable with children");
            const int32_t a = pod<int32_t>(), b = pod<int32_t>();
            nodes.emplace_back((int)a, (int)b);
        } else {
            if (!nc) fail("operation without operands");
            nodes.emplace_back((OpType)op, childs);
        }
    }
    const uint64_t r = pod<uint64_t>();
    if (r > left() / 4) fail("root list longer than the file");
    std::vector<Symbolic> roots;
    roots.reserve((size_t)r);
    for (uint64_t i = 0; i < r; ++i) {
        const uint32_t k = pod<uint32_t>();
        if (k >= nodes.size()) fail("root index out of range");
        roots.push_back(nodes[k]);
    }
    return roots;
}

}  // namespace Sym
