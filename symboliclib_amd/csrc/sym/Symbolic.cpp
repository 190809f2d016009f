#include "Symbolic.hpp"

#include <algorithm>
#include <cmath>

#include "Traverse.hpp"

namespace Sym {

namespace {

constexpr uint64_t splitmix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

constexpr hash_t opHash(int op) { return (hash_t)splitmix(0x53796d4f70000000ULL + (uint64_t)op); }

// Fixed-size node pool.  Nodes are recycled through an intrusive free list.
struct NodePool {
    std::vector<Symbolic::Node*> chunks;
    Symbolic::Node* freeList = nullptr;
    long long live = 0;
    static constexpr size_t kChunk = 1 << 16;

    Symbolic::Node* alloc() {
        if (!freeList) {
            auto* chunk = static_cast<Symbolic::Node*>(::operator new(sizeof(Symbolic::Node) * kChunk));
            chunks.push_back(chunk);
            for (size_t i = 0; i < kChunk; ++i) {
                chunk[i].ext = reinterpret_cast<Symbolic::Node**>(freeList);
                freeList = chunk + i;
            }
        }
        Symbolic::Node* n = freeList;
        freeList = reinterpret_cast<Symbolic::Node*>(n->ext);
        ++live;
        return n;
    }

    void free(Symbolic::Node* n) {
        n->ext = reinterpret_cast<Symbolic::Node**>(freeList);
        freeList = n;
        --live;
    }
};

NodePool& pool() {
    static NodePool* p = new NodePool();  // intentionally never destroyed
    return *p;
}

}  // namespace

const OpInfo OpInfos[NUM_OPS]{
    {255, false, false, false, "", "NOOP", opHash(0), 0},
    {0, false, false, true, "", "CONST", ConstHash, 0},
    {0, false, false, true, "", "VAR", VarHash, 1},
    {255, true, true, false, "+", "ADD", opHash(3), 1},
    {255, true, true, false, "*", "MUL", opHash(4), 1},
    {2, true, false, false, "*", "MULC", opHash(5), 0},
    {2, true, false, false, "/", "DIV", opHash(6), 1},
    {1, false, false, false, "sqrt", "SQRT", opHash(7), 5},
    {1, false, false, false, "sin", "SIN", opHash(8), 1},
    {1, false, false, false, "cos", "COS", opHash(9), 1},
    {1, false, false, false, "tan", "TAN", opHash(10), 1},
    {1, false, false, false, "cot", "COT", opHash(11), 1},
    {1, false, false, false, "asin", "ASIN", opHash(12), 1},
    {1, false, false, false, "acos", "ACOS", opHash(13), 1},
    {1, false, false, false, "atan", "ATAN", opHash(14), 1},
    {1, false, false, false, "exp", "EXP", opHash(15), 1},
    {1, false, false, false, "log", "LOG", opHash(16), 1},
    {2, false, false, false, "pow", "POW", opHash(17), 1},
    {2, true, false, false, "=", "ASSIGN", opHash(18), 1},
    {255, true, false, false, ";\n", "GROUP", opHash(19), 1},
    {3, false, false, false, "selectpos", "CONDPOS", opHash(20), 1},
    {3, false, false, false, "selectnonzero", "CONDNONZERO", opHash(21), 1},
    {3, false, false, false, "selectzero", "CONDZERO", opHash(22), 1},
    {1, false, false, false, "-", "NEG", opHash(23), 0},
    {2, false, false, false, "def", "DEFINE", opHash(24), 1},
    {255, true, false, false, ";\n", "BLOCK", opHash(25), 0},
    {1, false, false, false, "1./", "REC", opHash(26), 1},
    {2, false, false, false, "FIXED", "FIXED", opHash(27), 0},
    {0, false, false, false, "FUNC", "FUNC", opHash(28), 0},
    {255, false, false, false, "CALL", "CALL", opHash(29), 1},
    {255, false, false, false, "INSTANCE", "INSTANCE", opHash(30), 0},
};

int Symbolic::varIdCounter = 0;

long long Symbolic::liveNodes() { return pool().live; }

Symbolic Symbolic::generateUniqueVariable() { return Symbolic(varIdCounter++, EXPLICIT_INTERMEDIATE); }

Symbolic::Node* Symbolic::newNode(OpType op, uint32_t numChilds) {
    Node* n = pool().alloc();
    n->ref = 1;
    n->op = op;
    n->complexity = 0;
    n->spare = 0;
    n->numChilds = numChilds;
    n->tag = 0;
    n->algebraicHash = 0;
    n->inl[0] = n->inl[1] = nullptr;
    if (numChilds > 2) n->ext = new Node*[numChilds];
    return n;
}

void Symbolic::release(Node* root) {
    // iterative destruction so that long chains do not exhaust the call stack
    std::vector<Node*> stack{root};
    while (!stack.empty()) {
        Node* n = stack.back();
        stack.pop_back();
        if (n->op != CONST && n->op != VAR) {
            Node* const* ch = n->childArray();
            for (uint32_t i = 0; i < n->numChilds; ++i) {
                Node* c = ch[i];
                if (c && --c->ref == 0) stack.push_back(c);
            }
            if (n->numChilds > 2) delete[] n->ext;
        }
        pool().free(n);
    }
}

namespace {

inline hash_t wrapMul(hash_t a, hash_t b) { return (hash_t)((uint64_t)a * (uint64_t)b); }
inline hash_t wrapAdd(hash_t a, hash_t b) { return (hash_t)((uint64_t)a + (uint64_t)b); }

}  // namespace

void Symbolic::finishNode(Node* n) {
    const OpType op = n->op;
    hash_t h = 0;
    Node* const* ch = (op == CONST || op == VAR) ? nullptr : n->childArray();

    switch (op) {
        case MUL:
        case MULC:
            h = 1;
            for (uint32_t i = 0; i < n->numChilds; ++i) {
                if (ch[i]->op == ADD) h = wrapMul(h, hash(OpInfos[ADD].hash, ch[i]->algebraicHash));
                else h = wrapMul(h, ch[i]->algebraicHash);
            }
            break;
        case ADD:
            h = 0;
            for (uint32_t i = 0; i < n->numChilds; ++i) {
                if (ch[i]->op == MUL || ch[i]->op == MULC) h = wrapAdd(h, hash(OpInfos[MUL].hash, ch[i]->algebraicHash));
                else h = wrapAdd(h, ch[i]->algebraicHash);
            }
            break;
        case CONST: {
            const double c = n->constant;
            if (c == 0.0) h = 0;
            else if (std::fabs(c) < 9.2e18 && c == (double)(hash_t)c) h = (hash_t)c;
            else h = hash(c);
            break;
        }
        case VAR:
            h = hash(OpInfos[VAR].hash, hash(std::array<int, 2>{n->variable[0], n->variable[1]}));
            break;
        default: {
            if (op == DIV) {
                if (ch[0]->algebraicHash == 0) {
                    h = 0;
                    break;
                } else if (ch[0]->algebraicHash == ch[1]->algebraicHash) {
                    h = 1;
                    break;
                }
            }
            h = OpInfos[op].hash;
            if (op == INSTANCE) h = hash(h, (hash_t)n->tag);
            for (uint32_t i = 0; i < n->numChilds; ++i) {
                const OpType opi = ch[i]->op;
                if (opi == MULC || opi == MUL || opi == ADD) h = hash(h, hash(OpInfos[opi].hash, ch[i]->algebraicHash));
                else h = hash(h, ch[i]->algebraicHash);
            }
        }
    }
    n->algebraicHash = h;

    if (!n->numChilds || op == CONST || op == VAR) {
        n->complexity = OpInfos[op].cost;
    } else {
        unsigned tmp = 0;
        for (uint32_t i = 0; i < n->numChilds; ++i) tmp += ch[i]->complexity;
        n->complexity = (unsigned char)std::min(255u, n->numChilds * OpInfos[op].cost + tmp);
    }
}

Symbolic::Symbolic(double x) {
    n_ = newNode(CONST, 0);
    n_->constant = x;
    finishNode(n_);
}

Symbolic::Symbolic(int a, int b) {
    n_ = newNode(VAR, 0);
    n_->variable[0] = a;
    n_->variable[1] = b;
    finishNode(n_);
}

Symbolic::Symbolic(OpType op, const Symbolic& c0) {
    n_ = newNode(op, 1);
    n_->inl[0] = c0.n_;
    ++c0.n_->ref;
    finishNode(n_);
}

Symbolic::Symbolic(OpType op, const Symbolic& c0, const Symbolic& c1) {
    n_ = newNode(op, 2);
    n_->inl[0] = c0.n_;
    n_->inl[1] = c1.n_;
    ++c0.n_->ref;
    ++c1.n_->ref;
    finishNode(n_);
}

Symbolic::Symbolic(OpType op, const Symbolic& c0, const Symbolic& c1, const Symbolic& c2) {
    n_ = newNode(op, 3);
    n_->ext[0] = c0.n_;
    n_->ext[1] = c1.n_;
    n_->ext[2] = c2.n_;
    ++c0.n_->ref;
    ++c1.n_->ref;
    ++c2.n_->ref;
    finishNode(n_);
}

Symbolic::Symbolic(OpType op, const Symbolic* childs, unsigned numChilds) {
    n_ = newNode(op, numChilds);
    Node** dst = numChilds <= 2 ? n_->inl : n_->ext;
    for (unsigned i = 0; i < numChilds; ++i) {
        dst[i] = childs[i].n_;
        ++dst[i]->ref;
    }
    finishNode(n_);
}

Symbolic::Symbolic(OpType op, const std::vector<Symbolic>& childs) : Symbolic(op, childs.data(), (unsigned)childs.size()) {}

Symbolic Symbolic::instance(uint32_t templateId, const std::vector<Symbolic>& childs, hash_t) {
    Node* n = newNode(INSTANCE, (uint32_t)childs.size());
    n->tag = templateId;
    Node** dst = childs.size() <= 2 ? n->inl : n->ext;
    for (size_t i = 0; i < childs.size(); ++i) {
        dst[i] = childs[i].n_;
        ++dst[i]->ref;
    }
    finishNode(n);
    return Symbolic(n);
}

hash_t Symbolic::computeStructureHash() const {
    // hash of operation codes only: leaves of the same kind are interchangeable
    return traverseGenerate<hash_t>(*this, [](const Symbolic& x, std::vector<hash_t>& hashes) {
        hashes.push_back(OpInfos[x.op()].hash);
        if (x.op() == INSTANCE) hashes.push_back((hash_t)x.tag());
        return hashes.size() == 1 ? hashes.front() : hash(hashes);
    });
}


/* evaluation ****************************************************************/

double evaluate(const OpType id, const std::vector<double>& cvals) {
    double ret = -1.;
    switch (id) {
        case ADD:
            ret = 0.0;
            for (auto x : cvals) ret += x;
            break;
        case MULC:
            ret = cvals[0] * cvals[1];
            break;
        case MUL:
            ret = 1.;
            for (auto x : cvals) ret *= x;
            break;
        case DIV:
            ret = cvals[0] / cvals[1];
            break;
        case SQRT:
            ret = std::sqrt(cvals[0]);
            break;
        case NEG:
            ret = -cvals[0];
            break;
        case RECIPROCAL:
            ret = 1.0 / cvals[0];
            break;
        case POW:
            ret = std::pow(cvals[0], cvals[1]);
            break;
        default:
            std::cerr << "evaluate: operation " << OpInfos[id].code << " not supported\n";
            assert(0);
            break;
    }
    return ret;
}

/* operators (src/Operations.cpp:114-166) ************************************/

Symbolic operator*(const Symbolic& a, const Symbolic& b) {
    if (a.op() == CONST) return Symbolic(MULC, a, b);
    else if (b.op() == CONST) return Symbolic(MULC, b, a);
    else return Symbolic(MUL, a, b);
}

Symbolic operator*=(Symbolic& a, const Symbolic& b) {
    a = a * b;
    return a;
}

Symbolic operator+(const Symbolic& a, const Symbolic& b) { return Symbolic(ADD, a, b); }

Symbolic operator+=(Symbolic& a, const Symbolic& b) {
    a = Symbolic(ADD, a, b);
    return a;
}

Symbolic operator/(const Symbolic& a, const Symbolic& b) { return Symbolic(DIV, a, b); }

Symbolic operator/=(Symbolic& a, const Symbolic& b) {
    a = Symbolic(DIV, a, b);
    return a;
}

Symbolic operator-(const Symbolic& a, const Symbolic& b) { return Symbolic(ADD, a, Symbolic(MULC, Symbolic(-1.), b)); }

Symbolic operator-=(Symbolic& a, const Symbolic& b) {
    a = a - b;
    return a;
}

Symbolic operator-(const Symbolic& a) { return Symbolic(MULC, Symbolic(-1.), a); }

Symbolic sqrt(const Symbolic& a) { return Symbolic(SQRT, a); }

}  // namespace Sym
