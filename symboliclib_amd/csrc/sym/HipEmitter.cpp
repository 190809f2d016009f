// HIP source generation for a compiled Program (new emitter; the reference's HIP text is
// produced by src/ComputeUnit.cpp:368-450 with src/CodeGenerator.cpp:12-118).
//
// Each wave of independent kernels becomes one extern "C" __global__ function.  Its
// workgroups are split into contiguous ranges, one per kernel; a work-item evaluates one
// instance of its kernel's template in registers:
//   gathered inputs   x[p[base + i] + offset]   (structure-of-arrays index columns)
//   constants         c[base + i]
//   local temporaries const double rK = ...;
//   user outputs      out[o[base + i]] = ...    (final position, no gather pass)
//   intermediates     y[base + i] = ...         (read back by a later wave through p)
// All arithmetic is FP64.
//
// Every kernel takes x and y, which are the same value buffer (the runtime passes it twice).
// y, written, is not __restrict__; x keeps it because no launch reads through x what it
// writes through y: a wave reads only intermediates of earlier waves, and the split fused
// form's element launch writes the element images that only its sum launch reads.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>
#include <limits>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>

#include "Fused.hpp"
#include "Program.hpp"
#include "Tiling.hpp"
#include "Traverse.hpp"

namespace Sym {

namespace {

std::string literal(double d) {
    if (std::isnan(d)) return "__builtin_nan(\"\")";
    if (std::isinf(d)) return d > 0 ? "__builtin_huge_val()" : "(-__builtin_huge_val())";
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.17g", d);
    std::string s = buf;
    if (s.find_first_of(".eE") == std::string::npos) s += ".0";
    if (d < 0 || (d == 0 && std::signbit(d))) s = "(" + s + ")";
    return s;
}

std::string idx(size_t base, bool wide) {
    if (wide) return "(size_t)" + std::to_string(base) + "ull + i";
    return std::to_string(base) + "u + i";
}

struct KernelEmitter {
    const Program& prog;
    const ProgramKernel& K;
    bool wide;
    // tiled form: per-tile table D (operands | ids | outpos), cnt instances, LDS base L
    const TiledProgram* tiled = nullptr;
    size_t kernelIndex = 0;
    // fused form: element template evaluated into LDS (L[slot], L = the element's values)
    const FusedTemplate* fusedT = nullptr;

    std::string fusedLeaf(const std::array<int, 2>& v) const {
        switch (v[1]) {
            case -2:
                return "L[" + std::to_string(v[0]) + "]";
            case -3:
                return "a" + std::to_string(v[0]);
            case -4:
                return "q" + std::to_string(v[0]);
            case LOCAL_VAR:
                return "r" + std::to_string(v[0]);
            default:
                throw std::logic_error("emitHipFused: unexpected variable group " + std::to_string(v[1]));
        }
    }

    std::string tiledLeaf(const std::array<int, 2>& v) const {
        const int slot = v[0];
        const size_t V = K.numVarSlots;
        switch (v[1]) {
            case -2: {
                if (K.userOutputs) {
                    const size_t row = V + (K.numConstSlots ? 1 : 0) + (size_t)slot;
                    return "out[D[" + std::to_string(row) + "u * cnt + j]]";
                }
                return "lds[L + " + std::to_string(slot) + "u * cnt + j]";
            }
            case -3: {
                const auto& va = K.access.at((size_t)slot);
                const int kind = tiled->colKind[tiled->colOffset[kernelIndex] + va.column];
                const std::string e = "D[" + std::to_string(va.column) + "u * cnt + j]";
                const std::string off = va.offset ? ((va.offset > 0 ? " + " : " - ") + std::to_string(std::abs(va.offset)) + "u") : "";
                if (kind == 0) return "x[" + e + off + "]";
                if (kind == 1) return "lds[" + e + " & 0x7fffffffu]";
                return "gather(x, lds, " + e + ", " + std::to_string(va.offset) + ")";
            }
            case -4:
                return "c[" + std::to_string(K.constOffset + (size_t)slot * K.n) + "u + D[" + std::to_string(K.numVarSlots) +
                       "u * cnt + j]]";
            case LOCAL_VAR:
                return "r" + std::to_string(slot);
            default:
                throw std::logic_error("emitHipTiled: unexpected variable group " + std::to_string(v[1]));
        }
    }

    std::string leaf(const std::array<int, 2>& v) const {
        if (fusedT) return fusedLeaf(v);
        if (tiled) return tiledLeaf(v);
        const int slot = v[0];
        switch (v[1]) {
            case -2: {
                if (K.userOutputs) return "out[o[" + idx(K.outposOffset + (size_t)slot * K.n, wide) + "]]";
                return "y[" + idx(prog.numInputs + K.interBase + (size_t)slot * K.n, wide) + "]";
            }
            case -3: {
                const auto& va = K.access.at((size_t)slot);
                std::string s = "x[p[" + idx(K.indexOffset + (size_t)va.column * K.n, wide) + "]";
                if (va.offset) s += (va.offset > 0 ? " + " : " - ") + std::to_string(std::abs(va.offset)) + (wide ? "" : "u");
                return s + "]";
            }
            case -4:
                return "c[" + idx(K.constOffset + (size_t)slot * K.n, wide) + "]";
            case LOCAL_VAR:
                return "r" + std::to_string(slot);
            default:
                throw std::logic_error("emitHip: unexpected variable group " + std::to_string(v[1]));
        }
    }

    // expression text, built bottom-up without recursion
    std::string expr(const Symbolic& root) const {
        std::vector<std::string> st;
        postOrderTraverse(root, [&](const Symbolic& y) {
            const unsigned nc = y.numChilds();
            const OpType op = y.op();
            if (op == CONST) {
                st.push_back(literal(y.constant()));
                return;
            }
            if (op == VAR) {
                st.push_back(leaf(y.variable()));
                return;
            }
            std::vector<std::string> ch(st.end() - nc, st.end());
            st.erase(st.end() - nc, st.end());
            std::string s;
            auto join = [&](const char* sep) {
                s = "(";
                for (unsigned i = 0; i < nc; ++i) s += (i ? sep : "") + ch[i];
                s += ")";
            };
            switch (op) {
                case ADD:
                    join(" + ");
                    break;
                case MUL:
                case MULC:
                    join(" * ");
                    break;
                case DIV:
                    s = "(" + ch[0] + " / " + ch[1] + ")";
                    break;
                case NEG:
                    s = "(-" + ch[0] + ")";
                    break;
                case RECIPROCAL:
                    s = "(1.0 / " + ch[0] + ")";
                    break;
                case SQRT:
                    s = "sqrt(" + ch[0] + ")";
                    break;
                case POW:
                    s = "pow(" + ch[0] + ", " + ch[1] + ")";
                    break;
                case SIN:
                case COS:
                case TAN:
                case ASIN:
                case ACOS:
                case ATAN:
                case EXP:
                case LOG:
                    s = OpInfos[op].opSymbol + "(" + ch[0] + ")";
                    break;
                case COT:
                    s = "(1.0 / tan(" + ch[0] + "))";
                    break;
                default:
                    throw std::logic_error("emitHip: operation " + OpInfos[op].code + " cannot be emitted");
            }
            st.push_back(std::move(s));
        });
        return st.at(0);
    }

    void statements(const Symbolic& b, std::ostringstream& os, const std::string& ind) const {
        switch (b.op()) {
            case BLOCK:
                for (auto& c : b) statements(c, os, ind);
                break;
            case DEFINE:
                os << ind << "const double r" << b[0].variable()[0] << " = " << expr(b[1]) << ";\n";
                break;
            case ASSIGN: {
                if (b[0].op() != VAR) throw std::logic_error("emitHip: assignment to a non-variable");
                auto v = b[0].variable();
                if (v[1] == -2 && v[0] < 0) break;
                os << ind << leaf(v) << " = " << expr(b[1]) << ";\n";
                break;
            }
            default:
                throw std::logic_error("emitHip: unexpected statement " + OpInfos[b.op()].code);
        }
    }
};

// Straight-line template code as a statement list, for scheduling.  Leaves are keyed
// LOCAL k -> k, argument a -> ARG + a, parameter q -> PARAM + q.
struct Stmt {
    Symbolic s;              // DEFINE or ASSIGN
    int def = -1;            // LOCAL defined, or -1
    int out = -1;            // output slot assigned, or -1
    std::vector<int> uses;   // distinct leaf keys
};
constexpr int ARG = 1 << 24, PARAM = 1 << 25;

Symbolic withReciprocals(const Symbolic& y, const std::map<hash_t, int>& rcp) {
    if (y.numChilds() == 0) return y;
    if (y.op() == DIV) {
        auto it = rcp.find(y[1].ahash());
        if (it != rcp.end()) return Symbolic(MUL, withReciprocals(y[0], rcp), Symbolic(it->second, LOCAL_VAR));
    }
    std::vector<Symbolic> ch;
    for (auto& c : y) ch.push_back(withReciprocals(c, rcp));
    return Symbolic(y.op(), ch);
}

// A denominator divided by two or more times gets one reciprocal, defined once
// (x / d -> x * r, r = 1 / d): an FP64 division costs about ten multiplies on CDNA.
void shareReciprocals(std::vector<Symbolic>& seq) {
    std::map<hash_t, int> count;
    std::map<hash_t, Symbolic> denom;
    int nextLocal = 0;
    for (auto& b : seq) {
        if (b.op() == DEFINE) nextLocal = std::max(nextLocal, b[0].variable()[0] + 1);
        preOrderTraverse(b[1], [&](const Symbolic& y) {
            if (y.op() == VAR && y.variable()[1] == LOCAL_VAR) nextLocal = std::max(nextLocal, y.variable()[0] + 1);
            if (y.op() == DIV) {
                ++count[y[1].ahash()];
                denom.emplace(y[1].ahash(), y[1]);
            }
            return true;
        });
    }
    std::map<hash_t, int> rcp;
    for (auto& kv : count)
        if (kv.second >= 2) rcp[kv.first] = nextLocal++;
    if (rcp.empty()) return;
    // 1 / (u u) as (1 / u)^2 when 1 / u is shared too (a multiply instead of a division)
    std::map<hash_t, Symbolic> defOf;
    for (auto& b : seq)
        if (b.op() == DEFINE) defOf.emplace(b[0].ahash(), b[1]);
    std::vector<Symbolic> defs;
    for (auto& kv : rcp) {
        const Symbolic& d = denom.at(kv.first);
        Symbolic value(DIV, Symbolic(1.0), d);
        auto def = defOf.find(d.ahash());
        if (d.op() == VAR && def != defOf.end() && def->second.op() == MUL && def->second.numChilds() == 2 &&
            def->second[0].ahash() == def->second[1].ahash()) {
            auto u = rcp.find(def->second[0].ahash());
            if (u != rcp.end()) value = Symbolic(MUL, Symbolic(u->second, LOCAL_VAR), Symbolic(u->second, LOCAL_VAR));
        }
        defs.emplace_back(DEFINE, Symbolic(kv.second, LOCAL_VAR), value);
    }
    std::vector<Symbolic> out(defs.begin(), defs.end());
    for (auto& b : seq) out.emplace_back(b.op(), b[0], withReciprocals(b[1], rcp));
    seq.swap(out);
}

// Constant coefficients carried through a straight-line statement list.  Every value is kept
// as c * t with a numeric coefficient c and a "monic" expression t (t = 1 for constants):
// products multiply coefficients (negations flip them), quotients and reciprocals divide
// them, a sum (terms in canonical order) takes its smallest coefficient out of every term,
// sqrt and pow take theirs out when it is positive.  A definition whose monic expression equals an earlier one (or is
// a single variable) defines nothing new: it becomes an alias c * (earlier local).  Outputs
// multiply their coefficient back in.  Differentiated templates are full of x * 2, -x and
// 1 / (2 x) chains; carrying the constants removes most of those multiplies and the copies
// that only differ by a constant factor.  Coefficient products are exact for the small
// integer and power-of-two constants these chains carry; otherwise they reassociate a product
// of constants, a change within a few units in the last place.
class CoefficientFolding {
    struct CT {
        double c = 1.0;
        Symbolic t;  // empty: the value is the constant c
        bool one() const { return t.op() == NOOP; }
    };
    struct Alias {
        double c;
        Symbolic base;  // VAR of a defining local, argument or parameter; empty for a constant
    };
    std::map<int, Alias> local_;
    std::map<hash_t, Symbolic> byTerm_;  // monic expression -> local holding it
    std::map<hash_t, Symbolic> termOf_;  // local (VAR hash) -> its monic expression

    static Symbolic materialize(const CT& x) {
        if (x.one()) return Symbolic(x.c);
        if (x.c == 1.0) return x.t;
        if (x.c == -1.0) return Symbolic(NEG, x.t);
        return Symbolic(MUL, Symbolic(x.c), x.t);
    }
    static CT plain(const Symbolic& t) { return CT{1.0, t}; }

    CT norm(const Symbolic& y) {
        const OpType op = y.op();
        switch (op) {
            case CONST:
                return CT{y.constant(), Symbolic()};
            case VAR: {
                const auto v = y.variable();
                if (v[1] == LOCAL_VAR) {
                    auto it = local_.find(v[0]);
                    if (it != local_.end()) return CT{it->second.c, it->second.base};
                }
                return plain(y);
            }
            case NEG: {
                CT x = norm(y[0]);
                x.c = -x.c;
                return x;
            }
            case MUL:
            case MULC: {
                double c = 1.0;
                std::vector<Symbolic> ts;
                for (const auto& ch : y) {
                    CT x = norm(ch);
                    c *= x.c;
                    if (!x.one()) ts.push_back(x.t);
                }
                if (c == 0.0 || ts.empty()) return CT{c, Symbolic()};
                return CT{c, ts.size() == 1 ? ts[0] : Symbolic(MUL, ts)};
            }
            case DIV:
            case RECIPROCAL: {
                const CT a = op == DIV ? norm(y[0]) : CT{1.0, Symbolic()};
                const CT b = norm(op == DIV ? y[1] : y[0]);
                if (b.c == 0.0 || a.c == 0.0) return plain(Symbolic(DIV, materialize(a), materialize(b)));
                const double c = a.c / b.c;
                if (b.one()) return CT{c, a.t};
                // a / (u u) with an earlier 1 / u = r: a r r, multiplies instead of a division
                if (b.t.op() == VAR) {
                    auto def = termOf_.find(b.t.ahash());
                    if (def != termOf_.end() && def->second.op() == MUL && def->second.numChilds() == 2 &&
                        def->second[0].ahash() == def->second[1].ahash()) {
                        auto r = byTerm_.find(Symbolic(DIV, Symbolic(1.0), def->second[0]).ahash());
                        if (r != byTerm_.end()) {
                            const Symbolic rr(MUL, r->second, r->second);
                            return CT{c, a.one() ? rr : Symbolic(MUL, a.t, rr)};
                        }
                    }
                }
                return CT{c, Symbolic(DIV, a.one() ? Symbolic(1.0) : a.t, b.t)};
            }
            case ADD: {
                // combine equal monic terms, then factor the first term's coefficient out
                std::vector<CT> terms;
                std::map<hash_t, size_t> at;
                double k = 0.0;
                for (const auto& ch : y) {
                    CT x = norm(ch);
                    if (x.one()) {
                        k += x.c;
                        continue;
                    }
                    auto it = at.find(x.t.ahash());
                    if (it == at.end()) {
                        at.emplace(x.t.ahash(), terms.size());
                        terms.push_back(x);
                    } else {
                        terms[it->second].c += x.c;
                    }
                }
                std::vector<CT> nz;
                for (auto& x : terms)
                    if (x.c != 0.0) nz.push_back(x);
                if (nz.empty()) return CT{k, Symbolic()};
                // canonical term order, so a - b and b - a share one monic form
                std::stable_sort(nz.begin(), nz.end(), [](const CT& a, const CT& b) { return a.t.ahash() < b.t.ahash(); });
                // factor out the smallest coefficient (ties: the first), so the ratios left on
                // the other terms are at least 1 in magnitude (integers when the constants are)
                double c = nz[0].c;
                for (auto& x : nz)
                    if (std::fabs(x.c) < std::fabs(c)) c = x.c;
                std::vector<Symbolic> parts;
                for (auto& x : nz) parts.push_back(materialize(CT{x.c / c, x.t}));
                if (k != 0.0) parts.push_back(Symbolic(k / c));
                return CT{c, parts.size() == 1 ? parts[0] : Symbolic(ADD, parts)};
            }
            case SQRT: {
                const CT x = norm(y[0]);
                if (x.one()) return CT{std::sqrt(x.c), Symbolic()};
                if (x.c > 0.0) return CT{std::sqrt(x.c), Symbolic(SQRT, x.t)};
                return plain(Symbolic(SQRT, materialize(x)));
            }
            case POW: {
                const CT x = norm(y[0]), e = norm(y[1]);
                if (e.one() && x.c > 0.0 && !x.one()) return CT{std::pow(x.c, e.c), Symbolic(POW, x.t, Symbolic(e.c))};
                return plain(Symbolic(POW, materialize(x), materialize(e)));
            }
            default: {
                std::vector<Symbolic> ch;
                for (const auto& c : y) ch.push_back(materialize(norm(c)));
                return plain(Symbolic(op, ch));
            }
        }
    }

public:
    void run(std::vector<Symbolic>& seq) {
        std::vector<Symbolic> out;
        out.reserve(seq.size());
        for (const auto& b : seq) {
            const CT x = norm(b[1]);
            if (b.op() == ASSIGN) {
                out.emplace_back(ASSIGN, b[0], materialize(x));
                continue;
            }
            const int id = b[0].variable()[0];
            if (x.one() || x.t.op() == VAR) {
                local_[id] = Alias{x.c, x.t};
                continue;
            }
            auto it = byTerm_.find(x.t.ahash());
            if (it != byTerm_.end()) {
                local_[id] = Alias{x.c, it->second};
                continue;
            }
            const Symbolic v(id, LOCAL_VAR);
            byTerm_.emplace(x.t.ahash(), v);
            termOf_.emplace(v.ahash(), x.t);
            local_[id] = Alias{x.c, v};
            out.emplace_back(DEFINE, v, x.t);
        }
        seq.swap(out);
    }
};

// Products shared between statements: the pair of variable factors that occurs together in
// the most products (two or more) becomes one new local p = a * b, replacing a and b in each
// of those products; repeated until no pair repeats (greedy pair extraction).  Nested products
// inside one expression are flattened first.  Only variable factors (locals, arguments,
// parameters) pair; constants and compound factors stay where they are.
class ProductSharing {
    struct PNode {
        OpType op = NOOP;
        Symbolic leaf;                // CONST / VAR / other leaves
        std::vector<int> ch;          // children (PNode indices); for MUL the non-variable factors
        std::vector<Symbolic> vars;   // MUL only: variable factors
    };
    std::vector<PNode> nodes_;
    std::vector<int> muls_;

    int build(const Symbolic& y) {
        const OpType op = y.op();
        if (y.numChilds() == 0) {
            PNode n;
            n.op = op;
            n.leaf = y;
            nodes_.push_back(n);
            return (int)nodes_.size() - 1;
        }
        if (op == MUL || op == MULC) {
            std::vector<Symbolic> factors, st{y};
            while (!st.empty()) {
                Symbolic f = st.back();
                st.pop_back();
                if (f.op() == MUL || f.op() == MULC)
                    for (unsigned i = f.numChilds(); i-- > 0;) st.push_back(f[i]);
                else
                    factors.push_back(f);
            }
            PNode n;
            n.op = MUL;
            std::vector<int> ch;
            for (auto& f : factors) {
                if (f.op() == VAR) n.vars.push_back(f);
                else ch.push_back(build(f));
            }
            n.ch = ch;
            nodes_.push_back(n);
            muls_.push_back((int)nodes_.size() - 1);
            return (int)nodes_.size() - 1;
        }
        PNode n;
        n.op = op;
        std::vector<int> ch;
        for (const auto& c : y) ch.push_back(build(c));
        n.ch = ch;
        nodes_.push_back(n);
        return (int)nodes_.size() - 1;
    }

    Symbolic emit(int i) const {
        const PNode& n = nodes_[(size_t)i];
        if (n.op != MUL && n.ch.empty()) return n.leaf;
        std::vector<Symbolic> ch;
        if (n.op == MUL)
            for (auto& v : n.vars) ch.push_back(v);
        for (int c : n.ch) ch.push_back(emit(c));
        if (n.op == MUL && ch.size() == 1) return ch[0];
        return Symbolic(n.op, ch);
    }

public:
    void run(std::vector<Symbolic>& seq) {
        int nextLocal = 0;
        for (auto& b : seq) {
            if (b.op() == DEFINE) nextLocal = std::max(nextLocal, b[0].variable()[0] + 1);
            preOrderTraverse(b[1], [&](const Symbolic& y) {
                if (y.op() == VAR && y.variable()[1] == LOCAL_VAR) nextLocal = std::max(nextLocal, y.variable()[0] + 1);
                return true;
            });
        }
        std::vector<int> roots;
        for (auto& b : seq) roots.push_back(build(b[1]));
        std::vector<Symbolic> defs;
        for (;;) {
            std::map<std::pair<hash_t, hash_t>, int> count;
            std::map<hash_t, Symbolic> byHash;
            for (int m : muls_) {
                const auto& v = nodes_[(size_t)m].vars;
                std::set<std::pair<hash_t, hash_t>> seen;
                for (size_t i = 0; i < v.size(); ++i) {
                    byHash.emplace(v[i].ahash(), v[i]);
                    for (size_t j = i + 1; j < v.size(); ++j) {
                        hash_t a = v[i].ahash(), b = v[j].ahash();
                        if (b < a) std::swap(a, b);
                        if (seen.insert({a, b}).second) ++count[{a, b}];
                    }
                }
            }
            std::pair<hash_t, hash_t> best{};
            int bestCount = 1;
            for (auto& kv : count)
                if (kv.second > bestCount) {
                    bestCount = kv.second;
                    best = kv.first;
                }
            if (bestCount < 2) break;
            const Symbolic p(nextLocal++, LOCAL_VAR);
            defs.emplace_back(DEFINE, p, Symbolic(MUL, byHash.at(best.first), byHash.at(best.second)));
            for (int m : muls_) {
                auto& v = nodes_[(size_t)m].vars;
                int ia = -1, ib = -1;
                for (int i = 0; i < (int)v.size(); ++i) {
                    if (ia < 0 && v[(size_t)i].ahash() == best.first) ia = i;
                    else if (ib < 0 && v[(size_t)i].ahash() == best.second) ib = i;
                }
                if (ia < 0 || ib < 0) continue;
                v.erase(v.begin() + std::max(ia, ib));
                v.erase(v.begin() + std::min(ia, ib));
                v.push_back(p);
            }
        }
        if (defs.empty()) return;
        std::vector<Symbolic> out(defs.begin(), defs.end());
        for (size_t i = 0; i < seq.size(); ++i) out.emplace_back(seq[i].op(), seq[i][0], emit(roots[i]));
        seq.swap(out);
    }
};

std::vector<Stmt> flatten(const Symbolic& block, bool reciprocals = false, bool coefficients = false) {
    std::vector<Stmt> out;
    std::vector<Symbolic> seq;
    // BLOCKs nest only at the top; keep statement order
    std::function<void(const Symbolic&)> walk = [&](const Symbolic& b) {
        if (b.op() == BLOCK) {
            for (auto& c : b) walk(c);
            return;
        }
        seq.push_back(b);
    };
    walk(block);
    if (coefficients) CoefficientFolding().run(seq);
    if (reciprocals) shareReciprocals(seq);
    if (coefficients) ProductSharing().run(seq);
    for (auto& b : seq) {
        Stmt s;
        s.s = b;
        if (b.op() == DEFINE) s.def = b[0].variable()[0];
        else if (b.op() == ASSIGN) s.out = b[0].variable()[0];
        else throw std::logic_error("emitHipFused: unexpected statement " + OpInfos[b.op()].code);
        std::vector<int> u;
        preOrderTraverse(b[1], [&](const Symbolic& y) {
            if (y.op() == VAR) {
                const auto v = y.variable();
                if (v[1] == LOCAL_VAR) u.push_back(v[0]);
                else if (v[1] == -3) u.push_back(ARG + v[0]);
                else if (v[1] == -4) u.push_back(PARAM + v[0]);
            }
            return true;
        });
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        s.uses = std::move(u);
        out.push_back(std::move(s));
    }
    return out;
}

// Greedy list schedule of the statements in `keep` (a dependency-closed subset) that keeps
// few values live: among ready statements take the one that frees the most values and
// creates the fewest (outputs create none); ties keep the original order.
std::vector<size_t> lowPressureOrder(const std::vector<Stmt>& S, const std::vector<char>& keep) {
    const size_t n = S.size();
    std::map<int, size_t> defAt;
    for (size_t i = 0; i < n; ++i)
        if (keep[i] && S[i].def >= 0) defAt[S[i].def] = i;
    std::map<int, int> remaining;
    std::vector<int> indeg(n, 0);
    std::vector<std::vector<size_t>> users(n);
    for (size_t i = 0; i < n; ++i) {
        if (!keep[i]) continue;
        for (int u : S[i].uses) {
            ++remaining[u];
            auto it = defAt.find(u);
            if (u < ARG && it != defAt.end()) {
                ++indeg[i];
                users[it->second].push_back(i);
            }
        }
    }
    std::vector<size_t> ready, order;
    for (size_t i = 0; i < n; ++i)
        if (keep[i] && indeg[i] == 0) ready.push_back(i);
    std::set<int> live;
    while (!ready.empty()) {
        size_t bestPos = 0;
        long bestScore = std::numeric_limits<long>::min();
        for (size_t r = 0; r < ready.size(); ++r) {
            const Stmt& s = S[ready[r]];
            long score = 0;
            for (int u : s.uses) {
                if (remaining[u] == 1 && (u >= ARG ? true : live.count(u) > 0)) ++score;
                if (u >= ARG && !live.count(u) && remaining[u] > 1) --score;
            }
            if (s.def >= 0) --score;
            if (score > bestScore || (score == bestScore && ready[r] < ready[bestPos])) {
                bestScore = score;
                bestPos = r;
            }
        }
        const size_t i = ready[bestPos];
        ready.erase(ready.begin() + (long)bestPos);
        order.push_back(i);
        for (int u : S[i].uses) {
            if (--remaining[u] == 0) live.erase(u);
            else if (u >= ARG) live.insert(u);
        }
        if (S[i].def >= 0) live.insert(S[i].def);
        for (size_t j : users[i])
            if (--indeg[j] == 0) ready.push_back(j);
    }
    return order;
}

// Largest number of values live at once when the statements run in `order` (arguments count
// from their first use), and FP64 operation counts of those statements.
void scheduleStats(const std::vector<Stmt>& S, const std::vector<size_t>& order, FusedCodeStats& st) {
    std::map<int, size_t> last;
    for (size_t pos = 0; pos < order.size(); ++pos)
        for (int u : S[order[pos]].uses) last[u] = pos;
    std::set<int> live;
    size_t mx = 0;
    for (size_t pos = 0; pos < order.size(); ++pos) {
        const Stmt& s = S[order[pos]];
        for (int u : s.uses) live.insert(u);
        if (s.def >= 0 && last.count(s.def)) live.insert(s.def);
        mx = std::max(mx, live.size());
        for (auto it = live.begin(); it != live.end();)
            it = (last.count(*it) && last[*it] <= pos) ? live.erase(it) : std::next(it);
        preOrderTraverse(s.s[1], [&](const Symbolic& y) {
            const unsigned nc = y.numChilds();
            switch (y.op()) {
                case ADD: st.adds += nc - 1; break;
                case MUL:
                case MULC: st.muls += nc - 1; break;
                case DIV:
                case RECIPROCAL: ++st.divs; break;
                case SQRT: ++st.sqrts; break;
                default: break;
            }
            return true;
        });
        ++st.statements;
    }
    st.maxLive = std::max(st.maxLive, (int)mx);
}

}  // namespace

// CPU form of a program, after the reference's CPU path (src/ComputeUnit.cpp:409-437,
// src/StringResources.h:312-480): every kernel is a loop over its instances running the same
// straight-line template as the GPU waves form, split over OpenMP threads and left to the host
// compiler to vectorise; kernels of one wave share one parallel region (no barrier between
// them), waves run in order.  Entry point:
//   extern "C" void symlib_cpu_run(const double* x, double* y, const double* c,
//                                  const INDEX* p, const INDEX* o, double* out, int threads)
std::string emitCpu(const Program& prog) {
    const size_t maxIdx = std::max({prog.index.size(), prog.consts.size(), prog.outpos.size(), prog.numValues()});
    const bool wide = maxIdx >= (size_t(1) << 31);
    std::ostringstream os;
    os << "// generated by symboliclib_amd: CPU form of a program (OpenMP over instances), FP64\n"
       << "#include <math.h>\n"
       << "#include <stddef.h>\n"
       << "typedef unsigned int INDEX;\n\n";
    for (int w = 0; w < prog.numWaves; ++w) {
        os << "static void symlib_cpu_wave" << w
           << "(const double* __restrict__ x, double* y, const double* __restrict__ c, const INDEX* __restrict__ p,\n"
           << "    const INDEX* __restrict__ o, double* __restrict__ out, int threads) {\n"
           << "    (void)x; (void)y; (void)c; (void)p; (void)o; (void)out;\n"
           << "#pragma omp parallel num_threads(threads)\n"
           << "    {\n";
        for (const auto& K : prog.kernels) {
            if (K.wave != w || K.n == 0) continue;
            os << "#pragma omp for schedule(static) nowait\n"
               << "        for (" << (wide ? "size_t" : "unsigned") << " i = 0; i < " << K.n << (wide ? "ull" : "u") << "; ++i) {\n";
            KernelEmitter ke{prog, K, wide};
            const std::vector<Stmt> S = flatten(K.code, true, true);
            const std::vector<char> all(S.size(), 1);
            for (size_t si : lowPressureOrder(S, all)) ke.statements(S[si].s, os, "            ");
            os << "        }\n";
        }
        os << "    }\n"
           << "}\n\n";
    }
    os << "extern \"C\" void symlib_cpu_run(const double* x, double* y, const double* c, const INDEX* p, const INDEX* o,\n"
       << "    double* out, int threads) {\n";
    for (int w = 0; w < prog.numWaves; ++w) os << "    symlib_cpu_wave" << w << "(x, y, c, p, o, out, threads);\n";
    os << "}\n";
    return os.str();
}

std::string emitHip(const Program& prog, const EmitOptions& opt, std::vector<LaunchDesc>& launches) {
    const int T = opt.threadsPerBlock;
    if (T <= 0 || T > 1024 || T % 64) throw std::invalid_argument("emitHip: threads per block must be a multiple of 64 up to 1024");
    const size_t maxIdx = std::max({prog.index.size(), prog.consts.size(), prog.outpos.size(), prog.numValues()});
    const bool wide = maxIdx >= (size_t(1) << 31);

    std::ostringstream os;
    os << "// generated by symboliclib_amd: one function per wave of independent kernels, FP64\n";
    os << "typedef unsigned int INDEX;\n\n";

    launches.clear();
    for (int w = 0; w < prog.numWaves; ++w) {
        const std::string name = "symlib_wave" + std::to_string(w);
        os << "extern \"C\" __global__ void __launch_bounds__(" << T << ") " << name
           << "(const double* __restrict__ x, double* y, const double* __restrict__ c,\n"
           << "    const INDEX* __restrict__ p, const INDEX* __restrict__ o, double* __restrict__ out,\n"
           << "    const INDEX* __restrict__ tiles, unsigned* __restrict__ work) {\n";
        os << "    const unsigned b = blockIdx.x;\n";
        size_t blocks = 0;
        for (const auto& K : prog.kernels) {
            if (K.wave != w || K.n == 0) continue;
            const size_t nb = (K.n + T - 1) / T;
            const size_t b0 = blocks;
            blocks += nb;
            os << "    if (b < " << blocks << "u) {\n";
            os << "        const unsigned i = (b - " << b0 << "u) * " << T << "u + threadIdx.x;\n";
            os << "        if (i < " << K.n << "u) {\n";
            KernelEmitter ke{prog, K, wide};
            // the template's statements with constant coefficients folded, shared reciprocals
            // and products, in a low-register-pressure order
            const std::vector<Stmt> S = flatten(K.code, true, true);
            const std::vector<char> all(S.size(), 1);
            for (size_t si : lowPressureOrder(S, all)) ke.statements(S[si].s, os, "            ");
            os << "        }\n        return;\n    }\n";
        }
        os << "}\n\n";
        if (blocks > 0xffffffffull) throw std::length_error("emitHip: grid too large");
        launches.push_back({name, (uint32_t)blocks, (uint32_t)T, 0});
    }
    return os.str();
}

std::string emitHipTiled(const Program& prog, const TiledProgram& tp, const EmitOptions& opt,
                         std::vector<LaunchDesc>& launches) {
    if (!tp.valid) throw std::invalid_argument("emitHipTiled: program is not tiled");
    const int T = opt.threadsPerBlock;
    std::ostringstream os;
    os << "// generated by symboliclib_amd: tile-fused program, one workgroup per tile, FP64\n";
    os << "typedef unsigned int INDEX;\n\n";
    os << "__device__ __forceinline__ double gather(const double* __restrict__ x, const double* lds, INDEX v, int off) {\n"
       << "    return (v & 0x80000000u) ? lds[v & 0x7fffffffu] : x[(int)v + off];\n}\n\n";
    os << "extern \"C\" __global__ void __launch_bounds__(" << T << ") symlib_tiled"
       << "(const double* __restrict__ x, double* y, const double* __restrict__ c,\n"
       << "    const INDEX* __restrict__ p, const INDEX* __restrict__ o, double* __restrict__ out,\n"
       << "    const INDEX* __restrict__ tiles, unsigned* __restrict__ work) {\n";
    os << "    extern __shared__ double lds[];\n";
    os << "    const INDEX* __restrict__ H = tiles + tiles[blockIdx.x];\n";
    for (int w = 0; w < prog.numWaves; ++w) {
        std::vector<size_t> ks;
        for (size_t r = 0; r < prog.kernels.size(); ++r)
            if (prog.kernels[r].wave == w && tp.kernelUsed[r] && prog.kernels[r].n) ks.push_back(r);
        if (ks.empty()) continue;
        if (w > 0) os << "    __syncthreads();\n";
        os << "    {\n";
        std::string total;
        for (size_t q = 0; q < ks.size(); ++q) {
            os << "        const unsigned n" << q << " = H[" << 3 * ks[q] << "];\n";
            total += (q ? " + n" : "n") + std::to_string(q);
        }
        os << "        const unsigned total = " << total << ";\n";
        os << "        for (unsigned g = threadIdx.x; g < total; g += " << T << "u) {\n";
        os << "            unsigned j = g;\n";
        for (size_t q = 0; q < ks.size(); ++q) {
            const size_t r = ks[q];
            os << "            if (j < n" << q << ") {\n";
            os << "                const unsigned cnt = n" << q << ";\n";
            os << "                const INDEX* __restrict__ D = tiles + H[" << 3 * r + 1 << "];\n";
            if (!prog.kernels[r].userOutputs) os << "                const unsigned L = H[" << 3 * r + 2 << "];\n";
            KernelEmitter ke{prog, prog.kernels[r], false, &tp, r};
            ke.statements(prog.kernels[r].code, os, "                ");
            os << "                continue;\n            }\n";
            os << "            j -= n" << q << ";\n";
        }
        os << "        }\n    }\n";
    }
    os << "}\n";
    launches.clear();
    launches.push_back({"symlib_tiled", (uint32_t)tp.numTiles, (uint32_t)T, (uint32_t)tp.ldsBytes});
    return os.str();
}

// Fused assembly form (tile pool layout in Fused.hpp).  Per tile, three phases separated by
// barriers:
//   A0  the workgroup stages elements' arguments (gathered coordinates) and parameters, the
//       tile's recipe tables and its output constants into LDS.  With direct global-to-LDS
//       loads (gfx950 global_load_lds_dword: lane l of a wave writes dword l after an LDS base
//       held in M0) every copy is issued without a register round trip and without waiting,
//       and one wait precedes the barrier; a gathered double takes two lanes (low and high
//       dword).  Without them every thread loads into registers and stores to LDS;
//   A1  work items (element, output group) evaluate the element templates from LDS into the
//       tile's LDS values (groups split the outputs so a work item keeps few values live);
//   B   each wave takes 64 consecutive owned outputs, finds its lanes' contribution lists by
//       one wave prefix sum of the packed (count, constant flag) bytes, sums LDS values and
//       writes the result to its CSC position.
// Persistent schedules loop over tiles taken from an atomic counter in `work`; the last
// workgroup to finish clears the counter, so consecutive launches (and graph replays) start
// from zero.
namespace {

// Exclusive prefix sum over the 64 lanes of a wave with DPP moves (register crossbar, no LDS
// round trip): shifts by 1, 2, 4, 8 within each row of 16 lanes, then the row_bcast:15 and
// row_bcast:31 steps carry the sums of rows 0 and 0-1 into the rows above.  Lanes whose
// source is outside the row read 0 (bound_ctrl), lanes of masked-off rows keep `old` = 0.
void emitWaveScan(std::ostream& os) {
    os << "__device__ __forceinline__ unsigned symlib_excl_scan64(unsigned v, unsigned lane) {\n"
       << "    (void)lane;\n"
       << "    unsigned s = v;\n"
       << "    s += (unsigned)__builtin_amdgcn_mov_dpp((int)s, 0x111, 0xf, 0xf, true);\n"
       << "    s += (unsigned)__builtin_amdgcn_mov_dpp((int)s, 0x112, 0xf, 0xf, true);\n"
       << "    s += (unsigned)__builtin_amdgcn_mov_dpp((int)s, 0x114, 0xf, 0xf, true);\n"
       << "    s += (unsigned)__builtin_amdgcn_mov_dpp((int)s, 0x118, 0xf, 0xf, true);\n"
       << "    s += (unsigned)__builtin_amdgcn_update_dpp(0, (int)s, 0x142, 0xa, 0xf, false);\n"
       << "    s += (unsigned)__builtin_amdgcn_update_dpp(0, (int)s, 0x143, 0xc, 0xf, false);\n"
       << "    return s - v;\n}\n\n";
}

// Direct global-to-LDS copy helpers for a workgroup of T threads: symlib_lds_copy32 moves n
// dwords (64 per wave instruction); symlib_lds_gather64 moves doubles x[idx[i]], i < n, to
// dst[i] (32 per wave instruction, four index loads in flight per wave, then four direct
// loads).  Neither waits for the data; the caller waits (s_waitcnt 0) before its barrier.
void emitLdsDmaHelpers(std::ostream& os, int T) {
    os << "#define SYMLIB_LDS_DMA4(src, dst) __builtin_amdgcn_global_load_lds((const void*)(src), "
          "(__attribute__((address_space(3))) void*)(dst), 4, 0, 0)\n\n";
    // whole 64-word blocks unmasked, then at most one partial block; the source of a block is a
    // uniform base plus the lane's 32-bit byte offset, so the loop spends no vector instructions
    // on addresses (wv is uniform)
    os << "__device__ __forceinline__ void symlib_lds_copy32(const unsigned* __restrict__ src, unsigned* dst, unsigned n,\n"
       << "    unsigned wv, unsigned lane) {\n"
       << "    const unsigned off = lane << 2;\n"
       << "    unsigned b = wv * 64u;\n"
       << "    for (; b + 64u <= n; b += " << T << "u) SYMLIB_LDS_DMA4((const char*)(src + b) + off, dst + b);\n"
       << "    if (b < n && b + lane < n) SYMLIB_LDS_DMA4((const char*)(src + b) + off, dst + b);\n"
       << "}\n\n";
    os << "__device__ __forceinline__ void symlib_lds_gather64(const double* __restrict__ x, const INDEX* __restrict__ idx,\n"
       << "    double* dst, unsigned n, unsigned wv, unsigned lane) {\n"
       << "    for (unsigned b = wv * 128u; b < n; b += " << 2 * T << "u) {\n"
       << "        unsigned a[4];\n"
       << "#pragma unroll\n"
       << "        for (unsigned u = 0; u < 4u; ++u) {\n"
       << "            const unsigned i = b + 32u * u + (lane >> 1);\n"
       << "            a[u] = i < n ? idx[i] : 0u;\n"
       << "        }\n"
       // one wait for the four indices here: waited inside the lane-masked blocks below, each
       // would also wait for the direct loads issued before it
       << "        __asm__ volatile(\"\" :: \"v\"(a[0]), \"v\"(a[1]), \"v\"(a[2]), \"v\"(a[3]));\n"
       << "#pragma unroll\n"
       << "        for (unsigned u = 0; u < 4u; ++u) {\n"
       << "            const unsigned i = b + 32u * u + (lane >> 1);\n"
       << "            if (i < n) SYMLIB_LDS_DMA4((const unsigned*)(x + a[u]) + (lane & 1u), dst + b + 32u * u);\n"
       << "        }\n"
       << "    }\n"
       << "}\n\n";
}

// Product-sum tiles (buildProductTiles): A0 stages the tile's values of A (one gather per
// staged value, all independent) and its recipe in LDS; B sums, per output, A(i, j) and the products of LDS pairs.  Schedules and
// profiling stamps as in the element form.
std::string emitHipProducts(const FusedProgram& fp, const FusedOptions& opt, std::vector<LaunchDesc>& launches) {
    const int T = opt.threadsPerBlock;
    const bool persistent = opt.schedule != FusedSchedule::Static;
    const bool counter = persistent && opt.schedule != FusedSchedule::Strided;
    std::ostringstream os;
    auto stamp = [&](int k) {
        if (opt.profile) os << "        if (threadIdx.x == 0) prof[4u * it + " << k << "u] = wall_clock64();\n";
    };
    os << "// generated by symboliclib_amd: product-sum tiles (A^T A + A), FP64\n";
    os << "typedef unsigned int INDEX;\n\n";
    if (opt.ldsDma) emitLdsDmaHelpers(os, T);
    emitWaveScan(os);
    os << "extern \"C\" __global__ void __launch_bounds__(" << T;
    if (opt.wavesPerEU > 0) os << ", " << opt.wavesPerEU;
    os << ") symlib_fused"
       << "(const double* __restrict__ x, double* y, const double* __restrict__ c,\n"
       << "    const INDEX* __restrict__ p, const INDEX* __restrict__ o, double* __restrict__ out,\n"
       << "    const INDEX* __restrict__ tiles, unsigned* __restrict__ work) {\n";
    os << "    extern __shared__ double lds[];\n";
    if (counter) os << "    __shared__ unsigned s_next;\n";
    if (opt.profile) os << "    unsigned long long* __restrict__ prof = (unsigned long long*)(work + 4);\n";
    os << "    const unsigned numTiles = tiles[0];\n";
    // the wave index is uniform: loops over waves' shares run on scalar registers
    os << "    const unsigned lane = threadIdx.x & 63u, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);\n";
    os << "    for (unsigned it = blockIdx.x; it < numTiles;) {\n";
    stamp(0);
    os << "        const INDEX* __restrict__ H = tiles + tiles[1u + it];\n";
    // every header word first: read after a direct load to LDS, a header word becomes a vector
    // load whose wait also waits for every load issued before it
    os << "        const unsigned m = H[2], nch = H[3], words = H[10], nst = H[12];\n"
       << "        const unsigned hrec = H[4], hcnt = H[5], hslt = H[6], hchk = H[7], hres = H[9], hpos = H[13];\n";
    os << "        INDEX* __restrict__ R = (INDEX*)(lds + nst);\n";
    if (opt.ldsDma)
        os << "        symlib_lds_gather64(x, H + hpos, lds, nst, wv, lane);\n"
           << "        symlib_lds_copy32(H + hrec, R, words, wv, lane);\n"
           << "        __builtin_amdgcn_s_waitcnt(0);\n";
    else
        os << "        {\n"
           << "            const INDEX* __restrict__ PS = H + hpos;\n"
           << "            for (unsigned i = threadIdx.x; i < nst; i += " << T << "u) lds[i] = x[PS[i]];\n"
           << "            const INDEX* __restrict__ src = H + hrec;\n"
           << "            for (unsigned w = threadIdx.x; w < words; w += " << T << "u) R[w] = src[w];\n"
           << "        }\n";
    os << "        __syncthreads();\n";
    stamp(1);
    stamp(2);
    os << "        {\n"
       << "            const unsigned r0 = hrec;\n"
       << "            const unsigned char* __restrict__ C8 = (const unsigned char*)(R + (hcnt - r0));\n"
       << "            const unsigned short* __restrict__ S = (const unsigned short*)(R + (hslt - r0));\n"
       << "            const INDEX* __restrict__ CH = R + (hchk - r0);\n"
       << "            double* __restrict__ res = out + hres + R[0];\n"
       << "            for (unsigned ch = wv; ch < nch; ch += " << T / 64 << "u) {\n"
       << "                const unsigned oi = ch * 64u + lane;\n"
       << "                const bool act = oi < m;\n"
       << "                const unsigned b8 = act ? (unsigned)C8[oi] : 0u;\n"
       << "                const unsigned k = b8 & 0x7fu, lin = b8 >> 7;\n"
       << "                const unsigned base = CH[3u * ch] + symlib_excl_scan64(2u * k + lin, lane);\n"
       << "                double s = lin ? lds[S[base]] : 0.0;\n"
       << "                for (unsigned q = 0; q < k; ++q) s = fma(lds[S[base + lin + 2u * q]], lds[S[base + lin + 2u * q + 1u]], s);\n"
       << "                if (act) res[oi] = s;\n"
       << "            }\n"
       << "        }\n";
    os << "        __syncthreads();\n";
    stamp(3);
    if (counter)
        os << "        if (threadIdx.x == 0) s_next = gridDim.x + atomicAdd(work, 1u);\n"
           << "        __syncthreads();\n"
           << "        it = __builtin_amdgcn_readfirstlane(s_next);  // uniform: header loads stay scalar\n";
    else if (persistent)
        os << "        it += gridDim.x;\n";
    else
        os << "        break;\n";
    os << "    }\n";
    if (counter)
        os << "    if (threadIdx.x == 0) {\n"
           << "        __threadfence();\n"
           << "        if (atomicAdd(work + 1, 1u) == gridDim.x - 1u) {\n"
           << "            atomicExch(work, 0u);\n"
           << "            atomicExch(work + 1, 0u);\n"
           << "        }\n"
           << "    }\n";
    os << "}\n";
    launches.clear();
    if (!persistent && fp.numTiles > 0xffffffffull) throw std::length_error("emitHipFused: grid too large");
    const uint32_t grid = persistent ? LaunchDesc::GRID_FILL : (uint32_t)fp.numTiles;
    launches.push_back({"symlib_fused", grid, (uint32_t)T, (uint32_t)fp.ldsBytes});
    return os.str();
}

}  // namespace

std::string emitHipFused(const FusedProgram& fp, const FusedOptions& opt, std::vector<LaunchDesc>& launches,
                         FusedCodeStats* stats) {
    FusedCodeStats localStats;
    FusedCodeStats& cs = stats ? *stats : localStats;
    cs = FusedCodeStats();
    if (fp.kind == 1) return emitHipProducts(fp, opt, launches);
    const int T = opt.threadsPerBlock;
    if (T <= 0 || T > 1024 || T % 64) throw std::invalid_argument("emitHipFused: threads per block must be a multiple of 64 up to 1024");
    const bool persistent = opt.schedule != FusedSchedule::Static;
    const bool counter = persistent && opt.schedule != FusedSchedule::Strided;
    const bool blockSums = fp.sums == FusedSums::Blocks;
    static const Program noProgram;
    static const ProgramKernel noKernel;
    const size_t NT = fp.templates.size();
    auto stamp = [&](std::ostringstream& o, int k) {
        if (opt.profile)
            o << "        if (threadIdx.x == 0) prof[4u * it + " << k << "u] = wall_clock64();\n";
    };
    std::ostringstream os;
    os << "// generated by symboliclib_amd: fused assembly form (LDS-staged element tiles), FP64\n";
    os << "typedef unsigned int INDEX;\n";
    // two consecutive output slots: 16-byte aligned in LDS, 8-byte aligned in the results
    if (fp.sums == FusedSums::Blocks)
        os << "typedef double symlib_pair __attribute__((ext_vector_type(2)));\n"
           << "typedef double symlib_pair_u __attribute__((ext_vector_type(2), aligned(8)));\n";
    os << "\n";
    if (opt.ldsDma) emitLdsDmaHelpers(os, T);
    std::vector<int> groupsOf(NT);
    for (size_t t = 0; t < NT; ++t) {
        const FusedTemplate& ft = fp.templates[t];
        // A1: output groups, statements in a low-pressure order
        const std::vector<Stmt> S = flatten(ft.code, opt.shareReciprocals, opt.foldCoefficients);
        std::vector<size_t> outs;
        for (size_t i = 0; i < S.size(); ++i)
            if (S[i].out >= 0) outs.push_back(i);
        std::sort(outs.begin(), outs.end(), [&](size_t a, size_t b) { return S[a].out < S[b].out; });
        const int G = std::max(1, std::min<int>(opt.outputGroups, (int)outs.size()));
        groupsOf[t] = G;
        std::map<int, size_t> defAt;
        for (size_t i = 0; i < S.size(); ++i)
            if (S[i].def >= 0) defAt[S[i].def] = i;
        for (int g = 0; g < G; ++g) {
            const size_t o0 = outs.size() * (size_t)g / (size_t)G, o1 = outs.size() * (size_t)(g + 1) / (size_t)G;
            std::vector<char> keep(S.size(), 0);
            std::vector<size_t> st(outs.begin() + (long)o0, outs.begin() + (long)o1);
            while (!st.empty()) {
                const size_t i = st.back();
                st.pop_back();
                if (keep[i]) continue;
                keep[i] = 1;
                for (int u : S[i].uses)
                    if (u < ARG) {
                        auto it = defAt.find(u);
                        if (it != defAt.end()) st.push_back(it->second);
                    }
            }
            std::set<int> leaves;
            for (size_t i = 0; i < S.size(); ++i)
                if (keep[i])
                    for (int u : S[i].uses)
                        if (u >= ARG) leaves.insert(u);
            // leaves: arguments through the element's local node slots (E, numNodes columns of
            // cnt) into the tile's node coordinates (X, dof-major, nn per dof); parameters
            // from Q (columns of cnt)
            os << "__device__ __forceinline__ void symlib_elem" << t << "_" << g
               << "(const double* __restrict__ X, unsigned nn, const unsigned short* __restrict__ E,\n"
               << "    const double* __restrict__ Q, unsigned cnt, unsigned i, double* __restrict__ L) {\n"
               << "    (void)X; (void)nn; (void)E; (void)Q;\n";
            std::set<int> nodesUsed;
            for (int u : leaves)
                if (u < PARAM) nodesUsed.insert(ft.argNode[(size_t)(u - ARG)]);
            for (int j : nodesUsed) os << "    const unsigned e" << j << " = E[" << j << "u * cnt + i];\n";
            for (int u : leaves) {
                if (u < PARAM) {
                    const size_t a = (size_t)(u - ARG);
                    os << "    const double a" << a << " = X[" << ft.argDof[a] << "u * nn + e" << ft.argNode[a] << "];\n";
                } else {
                    os << "    const double q" << u - PARAM << " = Q[" << u - PARAM << "u * cnt + i];\n";
                }
            }
            KernelEmitter ke{noProgram, noKernel, false};
            ke.fusedT = &ft;
            const std::vector<size_t> order = lowPressureOrder(S, keep);
            for (size_t i : order) ke.statements(S[i].s, os, "    ");
            scheduleStats(S, order, cs);
            os << "}\n\n";
        }
    }
    emitWaveScan(os);
    // Per-template offsets inside a tile's staged region (parameter blocks, slot tables) and the
    // tile's value count, as expressions over its header words cnt<t> (emitTemplateWords).
    std::string vsum = "0u", psum = "0u", wsum = "0u";
    std::vector<std::string> pbase(NT), wbase(NT);
    for (size_t t = 0; t < NT; ++t) {
        const FusedTemplate& ft = fp.templates[t];
        const std::string cnt = "cnt" + std::to_string(t);
        pbase[t] = psum;
        wbase[t] = wsum;
        vsum += " + " + cnt + " * " + std::to_string(ft.valueStride) + "u";
        psum += " + " + cnt + " * " + std::to_string(ft.numParams) + "u";
        wsum += " + ((" + cnt + " * " + std::to_string(ft.numNodes) + "u + 1u) >> 1)";
    }
    // A0's node numbers sit in the header's table entry: every 32-node group a wave can take is
    // loaded up to the entry's capacity (zeros past the list) with the header words, so these
    // loads and the header's are one round trip
    const size_t nodeCap = fp.headerStride - fp.headerNodes;
    const size_t nodeRounds = (nodeCap + 32 * (size_t)(T / 64) - 1) / (32 * (size_t)(T / 64));
    auto emitNodeNumbers = [&](const std::string& ind, const std::string& H, const std::string& var, bool lanes) {
        for (size_t k = 0; k < nodeRounds; ++k) {
            if (lanes)
                os << ind << "const unsigned u" << k << " = wv + " << k * (size_t)(T / 64) << "u, s" << k << " = (u" << k
                   << " << 5) + (lane >> 1);\n";
            os << ind << "const unsigned " << var << k << " = s" << k << " < " << nodeCap << "u ? " << H << "["
               << fp.headerNodes << "u + s" << k << "] : 0u;\n";
        }
    };
    // Header words of the tile at H.  Every one is read before any direct load to LDS is issued:
    // read after one, a header word becomes a vector load whose wait also waits for every load
    // issued before it.
    auto emitTileWords = [&](const std::string& ind) {
        os << ind << "const unsigned B = H[0], N = H[1], m = H[2], nch = H[3], words = H[10];\n"
           << ind << "const unsigned hrec = H[4], hcnt = H[5], hslt = H[6], hchk = H[7], hcst = H[8], hres = H[9];\n"
           << ind << "const unsigned nconst = H[11] & 0x0fffffffu, kper = H[11] >> 28;  // stored constants, period\n";
    };
    auto emitTemplateWords = [&](const std::string& ind) {
        for (size_t t = 0; t < NT; ++t) {
            const size_t h = FUSED_HEADER + 3 * t;
            os << ind << "const unsigned cnt" << t << " = H[" << h << "], tab" << t << " = H[" << h + 1 << "], prm" << t << " = H["
               << h + 2 << "];\n";
        }
        os << ind << "const unsigned nn = H[13];\n";
        os << ind << "(void)B; (void)N; (void)m; (void)nch; (void)words; (void)nconst; (void)kper; (void)lane; (void)wv; (void)nn;\n"
           << ind << "(void)hrec; (void)hcnt; (void)hslt; (void)hchk; (void)hcst; (void)hres;\n";
        for (size_t t = 0; t < NT; ++t) os << ind << "(void)tab" << t << "; (void)prm" << t << ";\n";
        // staged region: parameters, node coordinates (stagedDofs per node), slot tables
        os << ind << "const unsigned np = " << psum << ", nsw = " << wsum << ";\n";
        os << ind << "const unsigned nvals = " << vsum << ", nstaged = np + " << fp.stagedDofs << "u * nn + ((nsw + 1u) >> 1);\n";
    };
    // A0 with direct loads into Ast / Xs / Es of the tile whose words are in scope
    auto emitStageDma = [&](bool earlyNodes, const std::string& nodeVar) {
        // A0 with direct loads: each wave takes 32 tile nodes at a time (two lanes per node:
        // low and high dword) and issues one direct load per staged dof; the parameter
        // columns (contiguous in c) and the slot tables (contiguous in the pool) are copied
        if (earlyNodes) {
            for (size_t k = 0; k < nodeRounds; ++k) {
                const std::string ks = std::to_string(k);
                os << "            if (s" << ks << " < nn) {\n"
                   << "                const unsigned* src = (const unsigned*)(x + B + " << nodeVar << ks << ") + (lane & 1u);\n";
                for (int d = 0; d < fp.stagedDofs; ++d) {
                    os << "                SYMLIB_LDS_DMA4(src";
                    if (d) os << " + " << 2 * d << "u * N";
                    os << ", Xs + " << d << "u * nn + (u" << ks << " << 5));\n";
                }
                os << "            }\n";
            }
        } else {
            os << "            const INDEX* __restrict__ tn = H + " << fp.headerNodes << "u;\n"
               << "            for (unsigned u = wv; u < ((nn + 31u) >> 5); u += " << T / 64 << "u) {\n"
               << "                const unsigned s = (u << 5) + (lane >> 1);\n"
               << "                if (s < nn) {\n"
               << "                    const unsigned* src = (const unsigned*)(x + B + tn[s]) + (lane & 1u);\n";
            for (int d = 0; d < fp.stagedDofs; ++d) {
                os << "                    SYMLIB_LDS_DMA4(src";
                if (d) os << " + " << 2 * d << "u * N";
                os << ", Xs + " << d << "u * nn + (u << 5));\n";
            }
            os << "                }\n"
               << "            }\n";
        }
        for (size_t t = 0; t < NT; ++t) {
            const FusedTemplate& ft = fp.templates[t];
            if (ft.numParams > 0)
                os << "            symlib_lds_copy32((const unsigned*)(c + prm" << t << "), (unsigned*)(Ast + " << pbase[t]
                   << "), " << 2 * ft.numParams << "u * cnt" << t << ", wv, lane);\n";
            os << "            symlib_lds_copy32(tiles + tab" << t << ", Es + " << wbase[t] << ", (cnt" << t << " * "
               << ft.numNodes << "u + 1u) >> 1, wv, lane);\n";
        }
    };
    // A1, element evaluation; copyRecipe: the recipe and output constants are copied to LDS by
    // direct loads issued first, overlapping the evaluation
    auto emitEval = [&](bool copyRecipe) {
        std::string total = "0u";
        for (size_t t = 0; t < NT; ++t) {
            os << "        const unsigned pc" << t << " = (cnt" << t << " + 63u) & ~63u;\n";
            total += " + " + std::to_string(groupsOf[t]) + "u * pc" + std::to_string(t);
        }
        os << "        {\n";
        if (copyRecipe)
            os << "            symlib_lds_copy32(tiles + hrec, R, words, wv, lane);\n";
        if (copyRecipe && !blockSums)
            os << "            if (nconst) symlib_lds_copy32((const unsigned*)(c + hcst), (unsigned*)K, 2u * nconst, wv, lane);\n";
        os << "            for (unsigned w = threadIdx.x; w < " << total << "; w += " << T << "u) {\n"
           << "                unsigned j = w, Lb = 0;\n";
        for (size_t t = 0; t < NT; ++t) {
            const FusedTemplate& ft = fp.templates[t];
            const std::string ts = std::to_string(t), cnt = "cnt" + ts, pc = "pc" + ts;
            os << "                if (j < " << groupsOf[t] << "u * " << pc << ") {\n"
               << "                    unsigned g = 0;\n";
            for (int g = 1; g < groupsOf[t]; ++g) os << "                    g += j >= " << g << "u * " << pc << " ? 1u : 0u;\n";
            os << "                    const unsigned i = j - g * " << pc << ";\n"
               << "                    if (i < " << cnt << ") {\n";
            for (int g = 0; g < groupsOf[t]; ++g)
                os << "                        " << (g ? "else " : "") << (g + 1 < groupsOf[t] ? "if (g == " + std::to_string(g) + "u) " : "")
                   << "symlib_elem" << t << "_" << g << "(Xs, nn, (const unsigned short*)(Es + " << wbase[t] << "), Ast + "
                   << pbase[t] << ", " << cnt << ", i, V + Lb + i * " << ft.valueStride << "u);\n";
            os << "                    }\n"
               << "                    continue;\n"
               << "                }\n"
               << "                j -= " << groupsOf[t] << "u * " << pc << ";\n"
               << "                Lb += " << cnt << " * " << ft.valueStride << "u;\n";
        }
        os << "            }\n"
           << "        }\n";
    };
    // B, sums
    auto emitSums = [&]() {
        os << "        {\n"
           << "            const unsigned r0 = hrec;\n"
           << "            const INDEX* __restrict__ Rn = R;\n"
           << "            const unsigned char* __restrict__ C8 = (const unsigned char*)(R + (hcnt - r0));\n"
           << "            const unsigned short* __restrict__ S = (const unsigned short*)(R + (hslt - r0));\n"
           << "            const INDEX* __restrict__ CH = R + (hchk - r0);\n"
           << "            double* __restrict__ res = out + hres;\n"
           << "            const unsigned long long below = lane == 63u ? ~0ull : ((2ull << lane) - 1ull);\n"
           // periodic constants: output block j of kper holds the stored block's constants again
           << "            const unsigned kblk = kper ? m / kper : 0u;\n"
           << "            const unsigned* __restrict__ S2 = (const unsigned*)S;  // slot pairs\n"
           ;
        if (fp.maxPairs <= FUSED_UNROLLED_PAIRS) {
            // SUM_CHUNKS chunks per wave and iteration, straight-line: every LDS read of the
            // chunks is issued before the first add (one chunk: the 256-cloth batch ran 1.5 %
            // faster than with two, whose extra waves' worth of work unbalances the last
            // round of chunks more than it hides latency).  A lane reads the program's
            // longest list length in pairs; words past its list (inside the slot table or the
            // chunk records after it) become the zero slot's pair, so the sums keep the list
            // order and the extra adds are + 0.0 (a sum that starts at +0.0 is never -0.0).
            // A chunk past the last one reads the first chunk's record and stores nothing.
            // Slots are byte offsets: an address is one mask or shift.
            constexpr int SUM_CHUNKS = 1;
            int maxPeriod = 0;
            for (uint32_t h : fp.tileHeader) maxPeriod = std::max(maxPeriod, (int)(fp.pool[h + 11] >> 28));
            os << "            const unsigned zp = (nvals << 3) | (nvals << 19);  // the zero slot's byte offset, twice\n"
               << "            for (unsigned ch0 = wv; ch0 < nch; ch0 += " << SUM_CHUNKS * (T / 64) << "u) {\n";
            for (int u = 0; u < SUM_CHUNKS; ++u) {
                const std::string x = "_" + std::to_string(u);
                if (u == 0)
                    os << "                const unsigned ch" << x << " = ch0;\n"
                       << "                const INDEX* __restrict__ cr" << x << " = CH + " << FUSED_CHUNK_WORDS << "u * ch0;\n";
                else
                    os << "                const unsigned ch" << x << " = ch0 + " << u * (T / 64) << "u;\n"
                       << "                const INDEX* __restrict__ cr" << x << " = CH + " << FUSED_CHUNK_WORDS << "u * (ch" << x
                       << " < nch ? ch" << x << " : ch0);\n";
                os << "                const unsigned oi" << x << " = ch" << x << " * 64u + lane;\n"
                   << "                const bool act" << x << " = oi" << x << " < m;\n"
                   << "                const unsigned b8" << x << " = act" << x << " ? (unsigned)C8[oi" << x << "] : 0u;\n"
                   << "                const unsigned k" << x << " = b8" << x << " & 0x7fu, hc" << x << " = b8" << x << " >> 7;\n"
                   << "                const unsigned ex" << x << " = symlib_excl_scan64(k" << x << " | (hc" << x << " << 16), lane);\n"
                   << "                const unsigned b2" << x << " = (cr" << x << "[0] + (ex" << x << " & 0xffffu)) >> 1, np2" << x
                   << " = k" << x << " >> 1;\n"
                   << "                const unsigned r" << x << " = cr" << x << "[1] + (unsigned)__popcll((((unsigned long long)cr" << x
                   << "[4] << 32) | cr" << x << "[3]) & below);\n"
                   << "                const unsigned pos" << x << " = Rn[2u * r" << x << "] + (oi" << x << " - Rn[2u * r" << x << " + 1u]);\n"
                   << "                unsigned kidx" << x << " = cr" << x << "[2] + (ex" << x << " >> 16);\n";
                if (maxPeriod > 1) {
                    // dof blocks that start at or before the chunk: one scalar correction;
                    // a block starting inside the chunk: a lane compare (chunk indices are
                    // uniform, the wave index being a scalar)
                    os << "                {\n"
                       << "                    unsigned j = 1, past = 0;\n"
                       << "                    for (; j < kper && j * kblk <= ch" << x << " * 64u; ++j) past += nconst;\n"
                       << "                    kidx" << x << " -= past;\n"
                       << "                    for (; j < kper && j * kblk < ch" << x << " * 64u + 64u; ++j) kidx" << x << " -= oi" << x
                       << " >= j * kblk ? nconst : 0u;\n"
                       << "                }\n";
                }
                // no constant: the zero slot
                os << "                const double kc" << x << " = *(hc" << x << " ? K + kidx" << x << " : V + nvals);\n";
                for (int q = 0; q < fp.maxPairs; ++q)
                    os << "                const unsigned a" << q << x << " = S2[b2" << x << " + " << q << "u], w" << q << x << " = " << q
                       << "u < np2" << x << " ? a" << q << x << " : zp;\n";
            }
            for (int u = 0; u < SUM_CHUNKS; ++u) {
                const std::string x = "_" + std::to_string(u);
                os << "                double s" << x << " = 0.0;\n";
                for (int q = 0; q < fp.maxPairs; ++q)
                    os << "                s" << x << " += *(const double*)((const char*)V + (w" << q << x << " & 0xffffu));\n"
                       << "                s" << x << " += *(const double*)((const char*)V + (w" << q << x << " >> 16));\n";
                // outputs are not read again by the launch: stored without caching (whole lines:
                // a wave's 64 consecutive outputs)
                os << "                if (act" << x << ") __builtin_nontemporal_store(s" << x << " + kc" << x << ", res + pos" << x
                   << ");\n";
            }
            os << "            }\n";
        } else {
            os << "            for (unsigned ch = wv; ch < nch; ch += " << T / 64 << "u) {\n"
               << "                const unsigned oi = ch * 64u + lane;\n"
               << "                const bool act = oi < m;\n"
               // the chunk record and the lane's count byte in one round of LDS reads
               << "                const INDEX* __restrict__ cr = CH + " << FUSED_CHUNK_WORDS << "u * ch;\n"
               << "                const unsigned b8 = act ? (unsigned)C8[oi] : 0u;\n"
               << "                const unsigned k = b8 & 0x7fu, hc = b8 >> 7;\n"
               << "                const unsigned ex = symlib_excl_scan64(k | (hc << 16), lane);\n"
               << "                const unsigned base = cr[0] + (ex & 0xffffu);  // even: lists have even length\n"
               // the lane's run: the chunk's first run plus the runs starting at lanes 1..lane
               << "                const unsigned r = cr[1] + (unsigned)__popcll((((unsigned long long)cr[4] << 32) | cr[3]) & below);\n"
               << "                const unsigned pos = Rn[2u * r] + (oi - Rn[2u * r + 1u]);\n"
               << "                unsigned kidx = cr[2] + (ex >> 16);\n"
               << "                for (unsigned j = 1; j < kper; ++j) kidx -= oi >= j * kblk ? nconst : 0u;\n"
               << "                double kc = 0.0;\n"
               << "                if (hc) kc = K[kidx];\n"
               << "                double s = 0.0;\n"
               << "                for (unsigned q = 0; q < k; q += 2u) {\n"
               << "                    const unsigned w = S2[(base + q) >> 1];\n"
               << "                    s += *(const double*)((const char*)V + (w & 0xffffu));\n"
               << "                    s += *(const double*)((const char*)V + (w >> 16));\n"
               << "                }\n"
               << "                if (act) res[pos] = hc ? s + kc : s;\n"
               << "            }\n";
        }
        os << "        }\n";
        };
    // B with block sums: B1 each lane sums one node block of the tile (sorted so a wave's
    // blocks have near-equal lists) from its list of (value index, transposed) entries into
    // registers; after a barrier B2 writes every block's entries to their tile outputs in LDS
    // (over the element values); after a second barrier B3 stores the outputs in CSC order
    auto emitBlockSums = [&]() {
        const int D = fp.blockDofs, Z = D * D, W = T / 64;
        const int iters = std::max<int>(1, (int)((fp.maxBlocks + (size_t)T - 1) / (size_t)T));
        // up to 8 blocks per thread: every block's sums in registers (unrolled); more (small
        // workgroups on large tiles): a loop over an indexed local array, which may live in
        // scratch memory
        const bool unrolled = iters <= 8;
        auto acc = [&](int it, int k) { return (unrolled ? "a" + std::to_string(it) : std::string("a")) + "_" + std::to_string(k); };
        auto recv = [&](int it) { return unrolled ? "r" + std::to_string(it) : std::string("rb"); };
        os << "        const unsigned nb = nch, noch = (m + 127u) >> 7, nbch = (nb + 63u) >> 6;\n"
           << "        const INDEX* __restrict__ BC = R + 5u * noch;  // block chunk records\n"
           << "        const INDEX* __restrict__ BR = BC + 2u * nbch;  // block records\n"
           << "        const INDEX* __restrict__ NR = R + (hchk - hrec);  // node records\n"
           << "        const INDEX* __restrict__ RN = R + (hcnt - hrec);  // runs\n"
           << "        const unsigned* __restrict__ LW = R + (hslt - hrec);  // lists\n"
           << "        const double* __restrict__ KG = c + (nconst ? hcst : 0u);  // block constants, in c[]\n";
        auto declare = [&](int it, const std::string& ind) {
            os << ind << "double";
            for (int k = 0; k < Z; ++k) os << (k ? ", " : " ") << acc(it, k) << " = 0.0";
            os << ";\n" << ind << "unsigned " << recv(it) << " = 0xc000u;  // no block\n";
        };
        if (unrolled)
            for (int it = 0; it < iters; ++it) declare(it, "        ");
        else
            os << "        double acc[" << iters << "][" << Z << "];\n        unsigned recs[" << iters << "];\n";
        // one entry of a diagonal block (D (D + 1) / 2 upper-triangle values) or of a full block
        // (D x D row-major, transposed when the element holds the pair the other way round)
        auto diagEntry = [&](int it, const std::string& e, const std::string& ind) {
            os << ind << "{\n" << ind << "    const double* __restrict__ p = V + (" << e << " & 0x7fffu);\n";
            int t = 0;
            for (int da = 0; da < D; ++da)
                for (int db = da; db < D; ++db, ++t) os << ind << "    " << acc(it, da * D + db) << " += p[" << t << "];\n";
            os << ind << "}\n";
        };
        auto fullEntry = [&](int it, const std::string& e, const std::string& ind) {
            os << ind << "{\n" << ind << "    const double* __restrict__ p = V + (" << e << " & 0x7fffu);\n"
               << ind << "    const bool tr = ((" << e << " >> 15) & 1u) != 0u;\n";
            for (int k = 0; k < Z; ++k) os << ind << "    const double v" << k << " = p[" << k << "];\n";
            for (int da = 0; da < D; ++da)
                for (int db = 0; db < D; ++db) {
                    const int k = da * D + db, kt = db * D + da;
                    if (k == kt) os << ind << "    " << acc(it, k) << " += v" << k << ";\n";
                    else os << ind << "    " << acc(it, k) << " += tr ? v" << kt << " : v" << k << ";\n";
                }
            os << ind << "}\n";
        };
        // B1 for iteration `it` (unrolled: a literal; looped: the loop's variable), in two parts:
        // b1pre reads the chunk's records, the lane's first list word and its block constant;
        // b1sum adds the entries.  Unrolled, every chunk's b1pre comes first, so the reads of a
        // wave's later chunks are in flight while its first chunk is summed.
        // Waves take the chunks of 64 blocks in snake order (every second pass reversed), so the
        // first, heaviest chunks (diagonal blocks, the longest lists) and the last, lightest ones
        // (blocks of a constant alone) fall to the same waves
        auto sfx = [&](int it) { return unrolled ? std::to_string(it) : std::string("b"); };
        auto b1pre = [&](int it) {
            const std::string rev = std::to_string(W - 1) + "u - wv", x = sfx(it);
            const std::string chx = unrolled ? std::to_string(it * W) + "u + " + (it % 2 ? "(" + rev + ")" : std::string("wv"))
                                             : "it * " + std::to_string(W) + "u + ((it & 1u) ? " + rev + " : wv)";
            os << "        const unsigned bch" << x << " = " << chx << ";\n"
               << "        unsigned bc" << x << " = 0u, bkw" << x << " = 0u, bw" << x << " = 0u;\n"
               << "        double bkc" << x << " = 0.0;  // the block's constant, added to its diagonal\n"
               << "        if (bch" << x << " < nbch) {  // uniform\n"
               << "            const unsigned bi = bch" << x << " * 64u + lane;\n"
               << "            const unsigned rec = bi < nb ? BR[bi] : 0xc000u;\n"
               << "            " << recv(it) << " = rec;\n"
               << "            bc" << x << " = BC[2u * bch" << x << "];\n"
               << "            bkw" << x << " = BC[2u * bch" << x << " + 1u];\n"
               << "            if ((bc" << x << " >> 24) != 0u) bw" << x << " = LW[(bc" << x << " & 0xffffffu) + lane];  // uniform\n"
            // constants are stored per block in block order, so a lane's is the chunk's first
            // plus the lanes below with one
               << "            const unsigned hc = (rec >> 30) & 1u;\n"
               << "            const unsigned long long cm = __ballot(hc);\n"
               << "            const unsigned ci = (bkw" << x << " & 0x7fffffffu) + __builtin_amdgcn_mbcnt_hi((unsigned)(cm >> 32), "
                  "__builtin_amdgcn_mbcnt_lo((unsigned)cm, 0u));\n"
               << "            if (hc) bkc" << x << " = KG[ci];\n"
               << "        }\n";
        };
        auto b1sum = [&](int it) {
            const std::string x = sfx(it);
            os << "        if (bch" << x << " < nbch) {  // uniform\n"
               << "            const unsigned L = bc" << x << " >> 24, kind = (" << recv(it) << " >> 14) & 3u;\n"
               << "            const unsigned* __restrict__ lw = LW + (bc" << x << " & 0xffffffu) + lane;\n";
            for (int kind = 0; kind < 2; ++kind) {
                os << "            " << (kind ? "else if" : "if") << " (kind == " << kind << "u) {\n"
                   << "                for (unsigned j = 0; j < L; j += 2u) {\n"
                   << "                    const unsigned w = j == 0u ? bw" << x << " : lw[(j >> 1) * 64u];\n";
                if (kind == 0) diagEntry(it, "w", "                    ");
                else fullEntry(it, "w", "                    ");
                os << "                    if (j + 1u < L) {\n"
                   << "                        const unsigned w1 = w >> 16;\n";
                if (kind == 0) diagEntry(it, "w1", "                        ");
                else fullEntry(it, "w1", "                        ");
                os << "                    }\n"
                   << "                }\n"
                   << "            }\n";
            }
            // a halved chunk (diagonal blocks with long lists): lane l + 32 summed the second half
            // of lane l's list; the halves meet in lane l, and lane l + 32 writes nothing (its
            // record has no constant)
            os << "            if (bkw" << x << " >> 31) {  // uniform\n";
            for (int da = 0; da < D; ++da)
                for (int db = da; db < D; ++db)
                    os << "                " << acc(it, da * D + db) << " += __shfl_xor(" << acc(it, da * D + db) << ", 32);\n";
            os << "                if (lane >= 32u) " << recv(it) << " = 0xc000u;\n"
               << "            }\n";
            for (int d = 0; d < D; ++d) os << "            " << acc(it, d * D + d) << " += bkc" << x << ";\n";
            if (D > 1) {
                os << "            if (kind == 0u) {  // the diagonal block's lower triangle\n";
                for (int da = 0; da < D; ++da)
                    for (int db = 0; db < da; ++db) os << "                " << acc(it, da * D + db) << " = " << acc(it, db * D + da) << ";\n";
                os << "            }\n";
            }
            os << "        }\n";
        };
        // B2 for iteration `it`: the block's entries to their tile outputs
        auto b2 = [&](int it) {
            const std::string r = recv(it);
            os << "        {\n"
               << "            const unsigned kind = (" << r << " >> 14) & 3u;\n"
               << "            if (kind < 3u) {\n"
               << "                const unsigned ni = " << r << " & 0x3fffu, rE = (" << r << " >> 16) & 0x7fu, rEK = (" << r
               << " >> 23) & 0x7fu;\n"
               << "                const unsigned w0 = NR[2u * ni], w1 = NR[2u * ni + 1u];\n"
               << "                const unsigned nE = (w1 >> 16) & 0xffu, nK = w1 >> 24;\n"
               << "                (void)nK;\n";
            const char* cs[3] = {"(w0 & 0xffffu)", "(w0 >> 16)", "(w1 & 0xffffu)"};
            for (int db = 0; db < D; ++db)
                for (int da = 0; da < D; ++da) {
                    std::string pos = std::string(cs[db]) + (da ? " + " + std::to_string(da) + "u * nE" : "") +
                                      (db < da ? " + nK" : "") + (da == db ? " + rEK" : " + rE");
                    if (da == db) os << "                V[" << pos << "] = " << acc(it, da * D + db) << ";\n";
                    else os << "                if (kind != 2u) V[" << pos << "] = " << acc(it, da * D + db) << ";\n";
                }
            os << "            }\n"
               << "        }\n";
        };
        if (unrolled) {
            for (int it = 0; it < iters; ++it) b1pre(it);
            for (int it = 0; it < iters; ++it) b1sum(it);
            os << "        __syncthreads();  // every element value read: the outputs overlay them\n";
            for (int it = 0; it < iters; ++it) b2(it);
        } else {
            os << "        for (unsigned it = 0; it < " << iters << "u; ++it) {\n";
            declare(0, "        ");
            b1pre(0);
            b1sum(0);
            for (int k = 0; k < Z; ++k) os << "        acc[it][" << k << "] = " << acc(0, k) << ";\n";
            os << "        recs[it] = " << recv(0) << ";\n"
               << "        }\n"
               << "        __syncthreads();  // every element value read: the outputs overlay them\n"
               << "        for (unsigned it = 0; it < " << iters << "u; ++it) {\n"
               << "        const double";
            for (int k = 0; k < Z; ++k) os << (k ? ", " : " ") << acc(0, k) << " = acc[it][" << k << "]";
            os << ";\n        const unsigned " << recv(0) << " = recs[it];\n";
            b2(0);
            os << "        }\n";
        }
        os << "        __syncthreads();\n";
        // B3: each lane stores a pair of consecutive output slots (16 bytes: one LDS read, one
        // store) and a wave 128 slots, whole lines, without caching; runs start at even slots,
        // so a pair never spans two runs, and a pair whose second slot is padding stores one
        os << "        {\n"
           << "            double* __restrict__ res = out + hres;\n"
           << "            const unsigned long long below = lane == 63u ? ~0ull : ((2ull << lane) - 1ull);\n"
           << "            for (unsigned ch = wv; ch < noch; ch += " << W << "u) {\n"
           << "                const unsigned s = ch * 128u + 2u * lane;\n"
           << "                const unsigned long long mk = ((unsigned long long)R[5u * ch + 2u] << 32) | R[5u * ch + 1u];\n"
           << "                const unsigned long long pd = ((unsigned long long)R[5u * ch + 4u] << 32) | R[5u * ch + 3u];\n"
           << "                const unsigned r = R[5u * ch] + (unsigned)__popcll(mk & below);\n"
           << "                const unsigned pos = RN[2u * r] + (s - RN[2u * r + 1u]);\n"
           << "                if (s < m) {\n"
           << "                    const symlib_pair v = *(const symlib_pair*)(V + s);\n"
           << "                    if ((pd >> lane) & 1ull) __builtin_nontemporal_store(v.x, res + pos);\n"
           << "                    else __builtin_nontemporal_store(v, (symlib_pair_u*)(res + pos));\n"
           << "                }\n"
           << "            }\n"
           << "        }\n";
    };
    // Kernel variants: 0 fused (stage, evaluate into LDS, sum), 1 split elements (stage, evaluate
    // into the tile's image in HBM), 2 split sums (copy image, recipe and constants to LDS, sum).
    auto kernel = [&](const char* name, int variant) {
        const bool stage = variant != 2, eval = variant != 2, sums = variant != 1;
        const unsigned ctr = variant == 2 ? 2u : 0u;  // work-queue counter words
        os << "extern \"C\" __global__ void __launch_bounds__(" << T;
        if (opt.wavesPerEU > 0) os << ", " << opt.wavesPerEU;
        os << ") " << name
           << "(const double* __restrict__ x, double* y, const double* __restrict__ c,\n"
           << "    const INDEX* __restrict__ p, const INDEX* __restrict__ o, double* __restrict__ out,\n"
           << "    const INDEX* __restrict__ tiles, unsigned* __restrict__ work) {\n";
        os << "    extern __shared__ double lds[];\n";
        if (counter) os << "    __shared__ unsigned s_next;\n";
        if (opt.profile) os << "    unsigned long long* __restrict__ prof = (unsigned long long*)(work + 4);\n";
        os << "    const unsigned numTiles = tiles[0];\n";
        // the wave index is uniform: loops over waves' shares run on scalar registers
    os << "    const unsigned lane = threadIdx.x & 63u, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);\n";
        os << "    for (unsigned it = blockIdx.x; it < numTiles;) {\n";
        if (variant != 2) stamp(os, 0);
        os << "        const INDEX* __restrict__ H = tiles + 1u + numTiles + it * " << fp.headerStride << "u;  // launch-order header\n";
        // A0's node numbers sit in the header's table entry: every 32-node group a wave can take
        // is loaded up to the entry's capacity (zeros past the list) first, so these loads and
        // the header's are one round trip
        const bool earlyNodes = stage && opt.ldsDma && nodeRounds <= 4;
        if (earlyNodes) emitNodeNumbers("        ", "H", "node", true);
        emitTileWords("        ");
        emitTemplateWords("        ");
        // V[nvals]: the tile's zero (odd contribution lists are padded with its slot)
        if (variant == 0 && blockSums)
            os << "        // values and their zeros, or the tile's outputs when more\n"
               << "        const unsigned vreg = nvals + " << fp.blockDofs * fp.blockDofs << "u > m ? nvals + " << fp.blockDofs * fp.blockDofs
               << "u : m;\n"
               << "        double* __restrict__ V = lds;\n        double* __restrict__ Ast = lds + vreg;\n";
        else if (variant == 0)
            os << "        double* __restrict__ V = lds;\n        double* __restrict__ Ast = lds + nvals + 1u;\n";
        if (variant == 1) os << "        double* __restrict__ V = y + H[12];\n        double* __restrict__ Ast = lds;\n";
        if (variant == 2 && blockSums)
            os << "        const unsigned vreg = nvals + " << fp.blockDofs * fp.blockDofs << "u > m ? nvals + " << fp.blockDofs * fp.blockDofs
               << "u : m;\n"
               << "        double* __restrict__ V = lds;\n        double* __restrict__ Ast = lds + vreg;\n";
        else if (variant == 2)
            os << "        double* __restrict__ V = lds;\n        double* __restrict__ Ast = lds + nvals + 1u;\n";
        os << "        INDEX* __restrict__ R = (INDEX*)(Ast + " << (variant == 2 ? "0u" : "nstaged") << ");\n";
        os << "        double* __restrict__ K = (double*)(R + ((words + 1u) & ~1u));\n";
        os << "        double* __restrict__ Xs = Ast + np;\n"
           << "        unsigned* __restrict__ Es = (unsigned*)(Xs + " << fp.stagedDofs << "u * nn);\n";
        os << "        (void)nvals; (void)nstaged; (void)R; (void)K; (void)V; (void)Xs; (void)Es;\n";
        os << "        {\n";
        if (sums && blockSums)
            os << "            if (threadIdx.x < " << fp.blockDofs * fp.blockDofs << "u) V[nvals + threadIdx.x] = 0.0;\n";
        else if (sums)
            os << "            if (threadIdx.x == 0u) V[nvals] = 0.0;\n";
        if (stage && opt.ldsDma) {
            emitStageDma(earlyNodes, "node");
        } else if (stage) {
            // A0 through registers: node coordinates, parameters and slot tables, one value per
            // thread and step
            os << "            const INDEX* __restrict__ tn = H + " << fp.headerNodes << "u;\n"
               << "            for (unsigned s = threadIdx.x; s < nn; s += " << T << "u) {\n"
               << "                const double* __restrict__ src = x + B + tn[s];\n";
            for (int d = 0; d < fp.stagedDofs; ++d)
                os << "                Xs[" << d << "u * nn + s] = src[" << d << "u * N];\n";
            os << "            }\n";
            for (size_t t = 0; t < NT; ++t) {
                const size_t h = FUSED_HEADER + 3 * t;
                const FusedTemplate& ft = fp.templates[t];
                if (ft.numParams > 0)
                    os << "            for (unsigned q = threadIdx.x; q < " << ft.numParams << "u * cnt" << t << "; q += " << T
                       << "u) Ast[" << pbase[t] << " + q] = c[H[" << h + 2 << "] + q];\n";
                os << "            for (unsigned q = threadIdx.x; q < ((cnt" << t << " * " << ft.numNodes << "u + 1u) >> 1); q += " << T
                   << "u) Es[" << wbase[t] << " + q] = tiles[H[" << h + 1 << "] + q];\n";
            }
        }
        if (opt.ldsDma) {
            if (variant == 2) os << "            symlib_lds_copy32((const unsigned*)(x + H[12]), (unsigned*)V, 2u * nvals, wv, lane);\n";
            // with element evaluation in the launch, the recipe and constants are copied during
            // it instead (A1), off the staging phase's critical path
            if (sums && !eval) os << "            symlib_lds_copy32(tiles + hrec, R, words, wv, lane);\n";
            if (sums && !eval && !blockSums)
                os << "            if (nconst) symlib_lds_copy32((const unsigned*)(c + hcst), (unsigned*)K, 2u * nconst, wv, lane);\n";
            os << "            __builtin_amdgcn_s_waitcnt(0);\n";
            // the node numbers stay live to here, so no register that held one is rewritten
            // while its load may be in flight (which would wait for every load issued so far)
            if (earlyNodes)
                for (size_t k = 0; k < nodeRounds; ++k) os << "            __asm__ volatile(\"\" :: \"v\"(node" << k << "));\n";
        } else {
            if (variant == 2)
                os << "            const double* __restrict__ img = x + H[12];\n"
                   << "            for (unsigned q = threadIdx.x; q < nvals; q += " << T << "u) V[q] = img[q];\n";
            if (sums)
                os << "            const INDEX* __restrict__ src = tiles + hrec;\n"
                   << "            for (unsigned w = threadIdx.x; w < words; w += " << T << "u) R[w] = src[w];\n";
            if (sums && !blockSums)
                os << "            if (nconst) {\n"
                   << "                const double* __restrict__ cs = c + hcst;\n"
                   << "                for (unsigned q = threadIdx.x; q < nconst; q += " << T << "u) K[q] = cs[q];\n"
                   << "            }\n";
        }
        os << "        }\n";
        os << "        __syncthreads();\n";
        if (variant != 2) stamp(os, 1);
        if (eval) {
            // A1: one loop over the (template, group, element) work items of the tile, elements
            // fastest; each (template, group) segment is padded to whole waves, so every wave
            // runs one element function (no divergent pair of them) and idle lanes sit at the
            // ends of segments
            emitEval(opt.ldsDma && sums);
            if (opt.ldsDma && sums) os << "        __builtin_amdgcn_s_waitcnt(0);\n";
            if (sums) os << "        __syncthreads();\n";
            stamp(os, 2);

        }
        if (sums && blockSums) emitBlockSums();
        else if (sums) emitSums();
        os << "        __syncthreads();\n";
        if (variant != 1) stamp(os, 3);
        if (counter) {
            os << "        if (threadIdx.x == 0) s_next = gridDim.x + atomicAdd(work + " << ctr << "u, 1u);\n"
               << "        __syncthreads();\n"
               << "        it = __builtin_amdgcn_readfirstlane(s_next);  // uniform: header loads stay scalar\n";
        } else if (persistent) {
            os << "        it += gridDim.x;\n";
        } else {
            os << "        break;\n";
        }
        os << "    }\n";
        if (counter)
            os << "    if (threadIdx.x == 0) {\n"
               << "        __threadfence();\n"
               << "        if (atomicAdd(work + " << ctr + 1 << "u, 1u) == gridDim.x - 1u) {\n"
               << "            atomicExch(work + " << ctr << "u, 0u);\n"
               << "            atomicExch(work + " << ctr + 1 << "u, 0u);\n"
               << "        }\n"
               << "    }\n";
        os << "}\n\n";
    };
    launches.clear();
    if (!persistent && fp.numTiles > 0xffffffffull) throw std::length_error("emitHipFused: grid too large");
    const uint32_t grid = persistent ? LaunchDesc::GRID_FILL : (uint32_t)fp.numTiles;
    if (opt.split) {
        kernel("symlib_fused_elem", 1);
        kernel("symlib_fused_sum", 2);
        launches.push_back({"symlib_fused_elem", grid, (uint32_t)T, (uint32_t)std::max<size_t>(8, fp.stagedLdsBytes)});
        launches.push_back({"symlib_fused_sum", grid, (uint32_t)T, (uint32_t)fp.ldsBytes});
    } else {
        kernel("symlib_fused", 0);
        launches.push_back({"symlib_fused", grid, (uint32_t)T, (uint32_t)fp.ldsBytes});
    }
    return os.str();
}

}  // namespace Sym
