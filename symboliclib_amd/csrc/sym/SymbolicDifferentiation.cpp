#include "SymbolicDifferentiation.hpp"

#include <cstring>
#include <stdexcept>

#include "Template.hpp"

namespace Sym {

double evaluateGroups(const Symbolic& x, const std::unordered_map<int, const double*>& groups) {
    return traverseGenerate<double>(x, [&](const Symbolic& y, std::vector<double>& childVals) {
        if (y.op() == CONST) return y.constant();
        if (y.op() == VAR) {
            auto v = y.variable();
            return groups.at(v[1])[v[0]];
        }
        return evaluate(y.op(), childVals);
    });
}

namespace {

// One reverse-mode step (src/Operations.cpp:10-60): push (child, adjoint) pairs.
void backprop(const Symbolic& expr, const Symbolic& d, std::vector<std::pair<Symbolic, Symbolic>>& stack) {
    switch (expr.op()) {
        case ASSIGN:
            stack.emplace_back(expr[1], d);
            break;
        case ADD:
            for (auto& c : expr) stack.emplace_back(c, d);
            break;
        case MUL:
            for (unsigned i = 0; i < expr.numChilds(); ++i) {
                std::vector<Symbolic> others;
                for (unsigned j = 0; j < expr.numChilds(); ++j)
                    if (j != i) others.push_back(expr[j]);
                others.push_back(d);
                stack.emplace_back(expr[i], Symbolic(MUL, others));
            }
            break;
        case MULC:
            stack.emplace_back(expr[1], d * expr[0]);
            break;
        case DIV:
            stack.emplace_back(expr[0], Symbolic(1.0) / expr[1] * d);
            stack.emplace_back(expr[1], -expr[0] / (expr[1] * expr[1]) * d);
            break;
        case RECIPROCAL:
            stack.emplace_back(expr[0], Symbolic(-1.) / (expr[0] * expr[0]) * d);
            break;
        case SQRT:
            stack.emplace_back(expr[0], Symbolic(0.5) / Symbolic(SQRT, expr[0]) * d);
            break;
        case NEG:
            stack.emplace_back(expr[0], -d);
            break;
        default:
            throw std::invalid_argument("differentiate: operation " + OpInfos[expr.op()].code + " not supported");
    }
}

inline bool nz(const Symbolic& s) { return s.ahash() != 0; }

// d/dvars of node y given the derivatives of its children (forward mode, shared factory).
std::vector<Symbolic> forwardRule(const Symbolic& y, std::vector<std::vector<Symbolic>>& cg,
                                  const std::unordered_map<hash_t, size_t>& varId, size_t nv, InternPool& pool,
                                  const Symbolic& zero) {
    std::vector<Symbolic> ret(nv, zero);
    switch (y.op()) {
        case ADD:
            for (size_t j = 0; j < nv; ++j) {
                std::vector<Symbolic> terms;
                for (auto& g : cg)
                    if (nz(g[j])) terms.push_back(g[j]);
                if (terms.size() == 1) ret[j] = terms[0];
                else if (terms.size() == 2) ret[j] = pool.make(ADD, terms[0], terms[1]);
                else if (terms.size() > 2) ret[j] = pool.make(ADD, terms);
            }
            break;
        case MULC:
            if (y[0].op() == CONST) {
                for (size_t j = 0; j < nv; ++j)
                    if (nz(cg[1][j])) ret[j] = pool.make(MULC, y[0], cg[1][j]);
                break;
            }
            // fallthrough: general product rule
        case MUL: {
            const unsigned n = y.numChilds();
            for (size_t j = 0; j < nv; ++j) {
                std::vector<Symbolic> terms;
                for (unsigned i = 0; i < n; ++i) {
                    if (!nz(cg[i][j])) continue;
                    if (n == 2) {
                        terms.push_back(pool.make(MUL, y[1 - i], cg[i][j]));
                    } else {
                        std::vector<Symbolic> f;
                        for (unsigned k = 0; k < n; ++k)
                            if (k != i) f.push_back(y[k]);
                        f.push_back(cg[i][j]);
                        terms.push_back(pool.make(MUL, f));
                    }
                }
                if (terms.size() == 1) ret[j] = terms[0];
                else if (terms.size() == 2) ret[j] = pool.make(ADD, terms[0], terms[1]);
                else if (terms.size() > 2) ret[j] = pool.make(ADD, terms);
            }
            break;
        }
        case DIV:
            for (size_t j = 0; j < nv; ++j) {
                const bool a = nz(cg[0][j]), b = nz(cg[1][j]);
                Symbolic t0, t1;
                if (a) t0 = pool.make(DIV, cg[0][j], y[1]);
                if (b) t1 = pool.make(MULC, Symbolic(-1.), pool.make(DIV, pool.make(MUL, y[0], cg[1][j]), pool.make(MUL, y[1], y[1])));
                if (a && b) ret[j] = pool.make(ADD, t0, t1);
                else if (a) ret[j] = t0;
                else if (b) ret[j] = t1;
            }
            break;
        case SQRT:
            for (size_t j = 0; j < nv; ++j)
                if (nz(cg[0][j])) ret[j] = pool.make(DIV, cg[0][j], pool.make(MULC, Symbolic(2.), y));
            break;
        case NEG:
            for (size_t j = 0; j < nv; ++j)
                if (nz(cg[0][j])) ret[j] = pool.make(MULC, Symbolic(-1.), cg[0][j]);
            break;
        case RECIPROCAL:
            for (size_t j = 0; j < nv; ++j)
                if (nz(cg[0][j])) ret[j] = pool.make(MULC, Symbolic(-1.), pool.make(DIV, cg[0][j], pool.make(MUL, y[0], y[0])));
            break;
        case VAR: {
            auto it = varId.find(y.ahash());
            if (it != varId.end()) ret[it->second] = Symbolic(1.);
            break;
        }
        case CONST:
            break;
        default:
            throw std::invalid_argument("hessianCached: operation " + OpInfos[y.op()].code + " not supported");
    }
    return ret;
}

}  // namespace

std::vector<Symbolic> differentiate(const Symbolic& expr, const std::vector<Symbolic>& vars) {
    const size_t numVars = vars.size();
    std::vector<Symbolic> grad(numVars, Symbolic(0.0));
    std::vector<std::pair<Symbolic, Symbolic>> stack{std::make_pair(expr, Symbolic(1.0))};
    std::unordered_map<hash_t, int> varMap;
    for (size_t i = 0; i < numVars; ++i) varMap[vars[i].ahash()] = (int)i;

    while (!stack.empty()) {
        auto x = stack.back();
        stack.pop_back();
        const OpType op = x.first.op();
        if (op == VAR) {
            auto it = varMap.find(x.first.ahash());
            if (it != varMap.end()) grad[it->second] += x.second;
        } else if (op != CONST) {
            backprop(x.first, x.second, stack);
        }
    }
    return grad;
}

std::vector<std::vector<Symbolic>> hessianCached(const Symbolic& x, const std::vector<Symbolic>& vars) {
    const size_t nv = vars.size();
    std::unordered_map<hash_t, size_t> varId;
    for (size_t i = 0; i < nv; ++i) varId[vars[i].ahash()] = i;
    InternPool pool;
    const Symbolic zero(0.0);
    auto diff = [&](const Symbolic& y, std::vector<std::vector<Symbolic>>& cg) {
        return forwardRule(y, cg, varId, nv, pool, zero);
    };
    std::unordered_map<long long, std::vector<Symbolic>, IdentityHash<long long>> cache;
    auto grad = traverseGenerateMap<std::vector<Symbolic>>(x, diff, cache, true);
    std::vector<std::vector<Symbolic>> hess(nv);
    for (size_t i = 0; i < nv; ++i) hess[i] = traverseGenerateMap<std::vector<Symbolic>>(grad[i], diff, cache, true);
    return hess;
}

std::vector<Symbolic> gradientCached(const Symbolic& x, const std::vector<Symbolic>& vars) {
    const size_t nv = vars.size();
    std::unordered_map<hash_t, size_t> varId;
    for (size_t i = 0; i < nv; ++i) varId[vars[i].ahash()] = i;
    InternPool pool;
    const Symbolic zero(0.0);
    std::unordered_map<long long, std::vector<Symbolic>, IdentityHash<long long>> cache;
    return traverseGenerateMap<std::vector<Symbolic>>(
        x, [&](const Symbolic& y, std::vector<std::vector<Symbolic>>& cg) { return forwardRule(y, cg, varId, nv, pool, zero); },
        cache, true);
}

bool isLiteralConstant(double c) { return c == (double)(long long)c && c > -1e9 && c < 1e9; }

void analyzeShape(const Symbolic& root, const std::unordered_map<hash_t, size_t>& varId, ElementShape& s) {
    analyzeShape(root, [&](const Symbolic& v) { return varId.count(v.ahash()) > 0; }, s);
}

void analyzeShape(const Symbolic& root, const std::function<bool(const Symbolic&)>& isDiff, ElementShape& s) {
    std::unordered_map<long long, int> num;
    std::unordered_map<hash_t, int> diffSlot;
    std::vector<std::pair<const Symbolic*, unsigned>> stack{{&root, 0u}};
    while (!stack.empty()) {
        const Symbolic* y = stack.back().first;
        unsigned& cnt = stack.back().second;
        if (cnt < y->numChilds()) {
            const Symbolic* c = y->begin() + cnt;
            ++cnt;
            if (!num.count(c->id())) stack.push_back({c, 0u});
            continue;
        }
        stack.pop_back();
        if (num.count(y->id())) continue;
        const int me = (int)s.nodes.size();
        num[y->id()] = me;
        s.nodes.push_back(*y);
        s.childIdx.emplace_back();
        const OpType op = y->op();
        if (op == VAR) {
            if (isDiff(*y)) {
                auto ds = diffSlot.find(y->ahash());
                int slot;
                if (ds == diffSlot.end()) {
                    slot = (int)s.diffLeaves.size();
                    diffSlot[y->ahash()] = slot;
                    s.diffLeaves.push_back(*y);
                } else slot = ds->second;
                s.leafKind.push_back(1);
                s.leafSlot.push_back(slot);
                s.sig.insert(s.sig.end(), {VAR, 1, slot});
            } else {
                s.leafKind.push_back(2);
                s.leafSlot.push_back((int)s.otherLeaves.size());
                s.sig.insert(s.sig.end(), {VAR, 2, (int64_t)s.otherLeaves.size()});
                s.otherLeaves.push_back(*y);
            }
        } else if (op == CONST) {
            if (isLiteralConstant(y->constant())) {
                s.leafKind.push_back(3);
                s.leafSlot.push_back(-1);
                int64_t bits;
                const double c = y->constant();
                std::memcpy(&bits, &c, sizeof(bits));
                s.sig.insert(s.sig.end(), {CONST, 3, bits});
            } else {
                s.leafKind.push_back(2);
                s.leafSlot.push_back((int)s.otherLeaves.size());
                s.sig.insert(s.sig.end(), {CONST, 2, (int64_t)s.otherLeaves.size()});
                s.otherLeaves.push_back(*y);
            }
        } else {
            s.leafKind.push_back(0);
            s.leafSlot.push_back(-1);
            s.sig.push_back(op);
            s.sig.push_back(y->numChilds());
            if (op == INSTANCE) s.sig.push_back(y->tag());
            for (auto& c : *y) {
                const int ci = num.at(c.id());
                s.childIdx[me].push_back(ci);
                s.sig.push_back(ci);
            }
        }
    }
    s.sig.push_back((int64_t)s.diffLeaves.size());
}

namespace {

// Differentiate a representative element once, over placeholder arguments.
const ExprTemplate& buildHessianTemplate(const ElementShape& s, hash_t sigHash) {
    const int nv = (int)s.diffLeaves.size();
    std::vector<Symbolic> ph(s.nodes.size());
    for (size_t k = 0; k < s.nodes.size(); ++k) {
        switch (s.leafKind[k]) {
            case 1:
                ph[k] = Symbolic(s.leafSlot[k], TEMPLATE_ARG);
                break;
            case 2:
                ph[k] = Symbolic(nv + s.leafSlot[k], TEMPLATE_ARG);
                break;
            case 3:
                ph[k] = s.nodes[k];
                break;
            default: {
                std::vector<Symbolic> ch;
                for (int ci : s.childIdx[k]) ch.push_back(ph[ci]);
                if (s.nodes[k].op() == INSTANCE) ph[k] = Symbolic::instance(s.nodes[k].tag(), ch, 0);
                else ph[k] = Symbolic(s.nodes[k].op(), ch);
            }
        }
    }
    std::vector<Symbolic> vars;
    for (int i = 0; i < nv; ++i) vars.emplace_back(i, TEMPLATE_ARG);
    auto h = hessianCached(ph.back(), vars);

    auto t = std::make_unique<ExprTemplate>();
    t->numArgs = nv + (int)s.otherLeaves.size();
    t->numDiffVars = nv;
    t->signature = s.sig;
    t->signatureHash = sigHash;
    // The Hessian is symmetric: entry (j, i) reuses the expression of (i, j), i <= j, so an
    // element computes each unordered pair once (the reference keeps both triangles in its
    // FIXED block although hessianSparse only reads one of them).
    t->entryToOutput.assign((size_t)nv * nv, -1);
    std::unordered_map<hash_t, int> unique;
    for (int i = 0; i < nv; ++i)
        for (int j = i; j < nv; ++j) {
            auto it = unique.find(h[i][j].ahash());
            int k;
            if (it == unique.end()) {
                k = (int)t->outputs.size();
                unique[h[i][j].ahash()] = k;
                t->outputs.push_back(h[i][j]);
            } else k = it->second;
            t->entryToOutput[(size_t)i * nv + j] = k;
            t->entryToOutput[(size_t)j * nv + i] = k;
        }
    t->numOutputs = (int)t->outputs.size();
    return getTemplate(registerTemplate(std::move(t)));
}

}  // namespace

const ExprTemplate& hessianTemplateFor(const ElementShape& s) {
    const hash_t sh = hash(s.sig);
    const ExprTemplate* t = findTemplate(sh, s.sig);
    return t ? *t : buildHessianTemplate(s, sh);
}

void sealBlock(std::vector<std::vector<Symbolic>>& entries) {
    std::unordered_map<hash_t, size_t> slotOf;  // algebraic hash -> index into assigns
    std::vector<Symbolic> assigns, vars;
    for (const auto& row : entries)
        for (const auto& e : row) {
            if (slotOf.count(e.ahash())) continue;
            slotOf.emplace(e.ahash(), assigns.size());
            vars.push_back(Symbolic::generateUniqueVariable());
            assigns.emplace_back(ASSIGN, vars.back(), e);
        }
    if (assigns.empty()) return;
    const Symbolic block(BLOCK, assigns);
    for (auto& row : entries)
        for (auto& e : row) e = Symbolic(FIXED, vars[slotOf.at(e.ahash())], block);
}

std::vector<std::tuple<int, int, Symbolic>> hessianSparseVars(const Symbolic& expr, const std::vector<Symbolic>& vars,
                                                              bool shareTemplates) {
    std::unordered_map<hash_t, size_t> varId;
    for (size_t i = 0; i < vars.size(); ++i) varId[vars[i].ahash()] = i;
    std::vector<std::tuple<int, int, Symbolic>> ret;
    if (expr.op() != ADD) throw std::invalid_argument("hessianSparse expects a sum of element terms");

    if (!shareTemplates) {
        for (auto& c : expr) {
            std::vector<Symbolic> vars2;
            std::vector<size_t> varIds;
            for (auto& v : harvestOp(VAR, c)) {
                auto it = varId.find(v.ahash());
                if (it != varId.end()) {
                    vars2.push_back(v);
                    varIds.push_back(it->second);
                }
            }
            auto h = hessianCached(c, vars2);
            sealBlock(h);
            for (size_t i = 0; i < varIds.size(); ++i)
                for (size_t j = 0; j < varIds.size(); ++j)
                    if (varIds[i] <= varIds[j]) ret.emplace_back((int)varIds[i], (int)varIds[j], h[i][j]);
        }
        return ret;
    }

    for (auto& c : expr) {
        ElementShape s;
        analyzeShape(c, varId, s);
        const ExprTemplate* t = &hessianTemplateFor(s);

        const int nv = (int)s.diffLeaves.size();
        std::vector<Symbolic> children;
        children.reserve(t->numOutputs + t->numArgs);
        for (int k = 0; k < t->numOutputs; ++k) children.push_back(Symbolic::generateUniqueVariable());
        for (auto& v : s.diffLeaves) children.push_back(v);
        for (auto& v : s.otherLeaves) children.push_back(v);
        Symbolic inst = Symbolic::instance(t->id, children, 0);

        std::vector<size_t> gid(nv);
        for (int i = 0; i < nv; ++i) gid[i] = varId.at(s.diffLeaves[i].ahash());
        for (int i = 0; i < nv; ++i)
            for (int j = 0; j < nv; ++j)
                if (gid[i] <= gid[j])
                    ret.emplace_back((int)gid[i], (int)gid[j],
                                     Symbolic(FIXED, children[t->entryToOutput[(size_t)i * nv + j]], inst));
    }
    return ret;
}

}  // namespace Sym
