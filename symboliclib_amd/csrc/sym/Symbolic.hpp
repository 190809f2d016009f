// Symbolic expression node: a reference-counted DAG node carrying an operation, its
// children (or a constant / variable payload), a commutativity-aware algebraic hash and a
// saturating complexity estimate.
//
// Semantics follow src/Symbolic.hpp:20-214 and src/Symbolic.cpp:61-131 of the reference:
//  * ADD hashes to the sum of its children's hashes and MUL/MULC to their product, so
//    a + b and b + a (or 2 * x and x + x) are recognised as the same expression;
//  * a CONST hashes to its value when the value is an integer, otherwise to the CRC of its
//    bytes; a VAR {id, group} hashes through the CRC of its two integers;
//  * complexity = numChilds * cost(op) + sum of child complexities, capped at 255.
// The node layout is new: 40-byte pooled nodes with up to two children stored inline.
// Nodes are not thread-safe (reference counts are plain integers), as in the reference.
#pragma once

#include <array>
#include <cassert>
#include <cstdint>
#include <iostream>
#include <unordered_map>
#include <vector>

#include "Operations.hpp"

namespace Sym {

constexpr hash_t VarHash = 0x6f916152447e9041LL;
constexpr hash_t ConstHash = 0x720980970d2b9f05LL;
constexpr hash_t NOOPHash = 0x2f2a32bb2a0f2e0dLL;

// Variable groups with a fixed meaning (src/Symbolic.hpp:13-15 of the reference).
constexpr int GLOBAL_INTERMEDIATE = 254;
constexpr int LOCAL_VAR = 253;
constexpr int EXPLICIT_INTERMEDIATE = 252;
// Groups used only inside expression templates (Template.hpp).
constexpr int TEMPLATE_ARG = 251;
constexpr int TEMPLATE_OUT = 250;

class Symbolic {
public:
    struct Node {
        uint32_t ref;
        OpType op;
        unsigned char complexity;
        uint16_t spare;
        uint32_t numChilds;
        uint32_t tag;  // INSTANCE: template id
        hash_t algebraicHash;
        union {
            Node* inl[2];
            Node** ext;
            double constant;
            int variable[2];
        };

        Node* const* childArray() const { return numChilds <= 2 ? inl : ext; }
    };

private:
    Node* n_ = nullptr;

    static Node* newNode(OpType op, uint32_t numChilds);
    static void finishNode(Node* n);
    static void release(Node* n);

    explicit Symbolic(Node* n) : n_(n) {}

public:
    static int varIdCounter;
    static long long liveNodes();

    // Fresh variable of group EXPLICIT_INTERMEDIATE (src/Symbolic.cpp:20-22).
    static Symbolic generateUniqueVariable();

    Symbolic() = default;
    Symbolic(double x);
    Symbolic(int a, int b);
    Symbolic(OpType op, const Symbolic& c0);
    Symbolic(OpType op, const Symbolic& c0, const Symbolic& c1);
    Symbolic(OpType op, const Symbolic& c0, const Symbolic& c1, const Symbolic& c2);
    Symbolic(OpType op, const std::vector<Symbolic>& childs);
    Symbolic(OpType op, const Symbolic* childs, unsigned numChilds);
    // INSTANCE node of template `templateId` with the given children.
    static Symbolic instance(uint32_t templateId, const std::vector<Symbolic>& childs, hash_t structureKey);

    Symbolic(const Symbolic& a) : n_(a.n_) {
        if (n_) ++n_->ref;
    }
    Symbolic(Symbolic&& a) noexcept : n_(a.n_) { a.n_ = nullptr; }
    Symbolic& operator=(const Symbolic& a) {
        if (a.n_) ++a.n_->ref;
        if (n_ && --n_->ref == 0) release(n_);
        n_ = a.n_;
        return *this;
    }
    Symbolic& operator=(Symbolic&& a) noexcept {
        if (this != &a) {
            if (n_ && --n_->ref == 0) release(n_);
            n_ = a.n_;
            a.n_ = nullptr;
        }
        return *this;
    }
    ~Symbolic() {
        if (n_ && --n_->ref == 0) release(n_);
    }

    hash_t computeStructureHash() const;

    inline OpType op() const { return n_ ? n_->op : NOOP; }
    inline std::array<int, 2> variable() const { return {n_->variable[0], n_->variable[1]}; }
    inline double constant() const { return n_->constant; }
    inline unsigned numChilds() const { return n_ ? n_->numChilds : 0; }
    inline uint32_t tag() const { return n_ ? n_->tag : 0; }

    inline const Symbolic& operator[](const size_t id) const {
        assert(n_ && id < n_->numChilds);
        return reinterpret_cast<const Symbolic*>(n_->childArray())[id];
    }
    inline const Symbolic* begin() const {
        return (n_ && n_->numChilds) ? reinterpret_cast<const Symbolic*>(n_->childArray()) : nullptr;
    }
    inline const Symbolic* end() const {
        return (n_ && n_->numChilds) ? reinterpret_cast<const Symbolic*>(n_->childArray()) + n_->numChilds
                                     : nullptr;
    }

    inline hash_t algebraicHash() const { return n_ ? n_->algebraicHash : NOOPHash; }
    inline hash_t ahash() const { return algebraicHash(); }
    inline unsigned char complexity() const { return n_ ? n_->complexity : 0; }

    inline bool operator<(const Symbolic& x) const { return ahash() < x.ahash(); }
    inline bool operator==(const Symbolic& x) const { return ahash() == x.ahash(); }

    // Pointer identity of the shared node.
    inline long long id() const { return (long long)(uintptr_t)n_; }
};

static_assert(sizeof(Symbolic) == sizeof(void*), "Symbolic must be a single pointer");

struct AlgebraicHashFunctor {
    inline hash_t operator()(const Symbolic& s) const { return s.ahash(); }
    inline bool operator()(const Symbolic& s0, const Symbolic& s1) const { return s0.ahash() < s1.ahash(); }
};

// Interning pool of the cached differentiation: an expression algebraically equal to one
// made through the pool before comes back as that earlier node, so equal derivative terms
// share one node, and one whose algebraic hash is a small integer k (the hash of the
// constant k) comes back as the constant k.
class InternPool {
    std::unordered_map<hash_t, Symbolic> nodes_;

public:
    template <class... Args>
    Symbolic make(const Args&... args) {
        const Symbolic built(args...);
        const hash_t h = built.ahash();
        if (isSmallConstant(h)) return Symbolic((double)h);
        return nodes_.emplace(h, built).first->second;
    }
};

}  // namespace Sym
