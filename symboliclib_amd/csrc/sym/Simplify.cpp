#include "Simplify.hpp"

#include <algorithm>
#include <cmath>
#include <unordered_map>
#include <vector>

#include "Traverse.hpp"

namespace Sym {

namespace {

bool foldable(OpType op) {
    switch (op) {
        case ADD:
        case MUL:
        case MULC:
        case DIV:
        case SQRT:
        case NEG:
        case RECIPROCAL:
        case POW:
            return true;
        default:
            return false;
    }
}

bool statement(OpType op) { return op == ASSIGN || op == DEFINE || op == BLOCK; }

// One node from its (already folded) operands: constants combined, all-constant operations
// evaluated, the small-integer hash rule applied.
Symbolic fold(const Symbolic& y, const std::vector<Symbolic>& ops) {
    const OpType op = y.op();
    Symbolic out;
    if (op == ADD || op == MUL || op == MULC) {
        const bool sum = op == ADD;
        double c = sum ? 0.0 : 1.0;
        std::vector<Symbolic> rest;
        for (const Symbolic& o : ops) {
            if (o.op() == CONST) c = sum ? c + o.constant() : c * o.constant();
            else rest.push_back(o);
        }
        if (!sum && c == 0.0) return Symbolic(0.0);
        if (c != (sum ? 0.0 : 1.0) || rest.empty()) rest.push_back(Symbolic(c));
        out = rest.size() == 1 ? rest.front() : Symbolic(sum ? ADD : MUL, rest);
    } else if (foldable(op) && std::all_of(ops.begin(), ops.end(), [](const Symbolic& o) { return o.op() == CONST; })) {
        std::vector<double> v;
        for (const Symbolic& o : ops) v.push_back(o.constant());
        out = Symbolic(evaluate(op, v));
    } else {
        out = op == INSTANCE ? Symbolic::instance(y.tag(), ops, 0) : Symbolic(op, ops);
    }
    if (!statement(out.op()) && out.op() != CONST && isSmallConstant(out.ahash())) return Symbolic((double)out.ahash());
    return out;
}

// ---- products

struct Power {
    Symbolic base;
    double exponent;
};

// Add the powers of the factors of x, raised to `e`, into `pw` (by algebraic hash of the base,
// in order of first appearance); constants multiply `coef`.
void collectPowers(const Symbolic& x, double e, std::vector<Power>& pw, std::unordered_map<hash_t, size_t>& at,
                   double& coef) {
    switch (x.op()) {
        case CONST:
            coef *= std::pow(x.constant(), e);
            return;
        case MUL:
        case MULC:
            for (const Symbolic& c : x) collectPowers(c, e, pw, at, coef);
            return;
        case DIV:
            collectPowers(x[0], e, pw, at, coef);
            collectPowers(x[1], -e, pw, at, coef);
            return;
        case SQRT:
            collectPowers(x[0], 0.5 * e, pw, at, coef);
            return;
        case POW:
            if (x[1].op() == CONST) {
                collectPowers(x[0], e * x[1].constant(), pw, at, coef);
                return;
            }
            break;
        default:
            break;
    }
    auto it = at.find(x.ahash());
    if (it == at.end()) {
        at.emplace(x.ahash(), pw.size());
        pw.push_back({x, e});
    } else {
        pw[it->second].exponent += e;
    }
}

Symbolic product(std::vector<Symbolic> f) {
    if (f.empty()) return Symbolic(1.0);
    if (f.size() == 1) return f.front();
    return Symbolic(MUL, f);
}

Symbolic normalizeProduct(const Symbolic& x) {
    std::vector<Power> pw;
    std::unordered_map<hash_t, size_t> at;
    double coef = 1.0;
    collectPowers(x, 1.0, pw, at, coef);
    if (coef == 0.0) return Symbolic(0.0);
    std::vector<Symbolic> num, den, numHalf, denHalf, other;
    for (const Power& p : pw) {
        const double e = std::fabs(p.exponent);
        if (e == 0.0) continue;
        const double whole = std::floor(e);
        const bool half = e - whole == 0.5;
        if (e != whole && !half) {
            other.emplace_back(POW, p.base, Symbolic(p.exponent));
            continue;
        }
        auto& full = p.exponent > 0 ? num : den;
        for (int k = 0; k < (int)whole; ++k) full.push_back(p.base);
        if (half) (p.exponent > 0 ? numHalf : denHalf).push_back(p.base);
    }
    num.insert(num.end(), other.begin(), other.end());
    if (!numHalf.empty()) {
        const Symbolic r = denHalf.empty() ? product(numHalf) : Symbolic(DIV, product(numHalf), product(denHalf));
        num.emplace_back(SQRT, r);
    } else if (!denHalf.empty()) {
        den.emplace_back(SQRT, product(denHalf));
    }
    if (coef != 1.0) num.push_back(Symbolic(coef));
    Symbolic n = product(num);
    if (den.empty()) return n;
    return Symbolic(DIV, n, product(den));
}

// ---- sums

struct Term {
    double coef;
    Symbolic t;
};

void addTerm(const Symbolic& x, double m, std::vector<Term>& terms, std::unordered_map<hash_t, size_t>& at) {
    auto it = at.find(x.ahash());
    if (it == at.end()) {
        at.emplace(x.ahash(), terms.size());
        terms.push_back({m, x});
    } else {
        terms[it->second].coef += m;
    }
}

// Collect the terms of x times `m` (nested sums, negations, constant multiples and the
// constant factors of products), merging algebraically equal terms.  A constant multiple of a
// sum stays one term (it is not distributed).
void collectTerms(const Symbolic& x, double m, std::vector<Term>& terms, std::unordered_map<hash_t, size_t>& at,
                  double& constant) {
    switch (x.op()) {
        case CONST:
            constant += m * x.constant();
            return;
        case ADD:
            for (const Symbolic& c : x) collectTerms(c, m, terms, at, constant);
            return;
        case NEG:
            collectTerms(x[0], -m, terms, at, constant);
            return;
        case MULC:
        case MUL: {
            double k = 1.0;
            std::vector<Symbolic> rest;
            for (const Symbolic& c : x) {
                if (c.op() == CONST) k *= c.constant();
                else rest.push_back(c);
            }
            if (rest.empty()) {
                constant += m * k;
                return;
            }
            if (k != 1.0) {
                const Symbolic r = product(rest);
                if (r.op() == ADD) addTerm(r, m * k, terms, at);
                else collectTerms(r, m * k, terms, at, constant);
                return;
            }
            break;
        }
        default:
            break;
    }
    addTerm(x, m, terms, at);
}

std::vector<Symbolic> factorsOf(const Symbolic& t) {
    if (t.op() == MUL) return std::vector<Symbolic>(t.begin(), t.end());
    return {t};
}

Symbolic scaled(double c, const Symbolic& t) {
    if (c == 1.0) return t;
    return Symbolic(MULC, Symbolic(c), t);
}

Symbolic buildSum(std::vector<Term> terms, double constant);

// Factor out, while one is shared by two or more terms, the factor shared by the most terms
// (then the most complex); the rest of the terms, and the constant, are added after.
Symbolic factorSum(const std::vector<Term>& terms, double constant) {
    std::vector<std::vector<Symbolic>> fs;
    std::unordered_map<hash_t, std::vector<size_t>> holders;
    std::vector<hash_t> order;
    for (size_t i = 0; i < terms.size(); ++i) {
        fs.push_back(factorsOf(terms[i].t));
        std::vector<hash_t> seenHere;
        for (const Symbolic& f : fs.back()) {
            const hash_t h = f.ahash();
            if (std::find(seenHere.begin(), seenHere.end(), h) != seenHere.end()) continue;
            seenHere.push_back(h);
            auto& hs = holders[h];
            if (hs.empty()) order.push_back(h);
            hs.push_back(i);
        }
    }
    hash_t best = 0;
    size_t bestCount = 1;
    int bestComplexity = -1;
    Symbolic bestFactor;
    for (hash_t h : order) {
        const auto& hs = holders[h];
        if (hs.size() < 2) continue;
        const Symbolic& f = *std::find_if(fs[hs.front()].begin(), fs[hs.front()].end(),
                                          [&](const Symbolic& s) { return s.ahash() == h; });
        if (hs.size() > bestCount || (hs.size() == bestCount && f.complexity() > bestComplexity)) {
            best = h;
            bestCount = hs.size();
            bestComplexity = f.complexity();
            bestFactor = f;
        }
    }
    if (bestCount < 2) return Symbolic();  // nothing to factor
    std::vector<Term> inner, outer;
    const auto& hs = holders[best];
    size_t k = 0;
    for (size_t i = 0; i < terms.size(); ++i) {
        if (k < hs.size() && hs[k] == i) {
            ++k;
            std::vector<Symbolic> rest = fs[i];
            rest.erase(std::find_if(rest.begin(), rest.end(), [&](const Symbolic& s) { return s.ahash() == best; }));
            inner.push_back({terms[i].coef, product(rest)});
        } else {
            outer.push_back(terms[i]);
        }
    }
    const Symbolic factored = Symbolic(MUL, bestFactor, buildSum(inner, 0.0));
    outer.push_back({1.0, factored});
    return buildSum(outer, constant);
}

Symbolic buildSum(std::vector<Term> terms, double constant) {
    for (const Term& t : terms)
        if (t.t.op() == CONST) constant += t.coef * t.t.constant();
    terms.erase(std::remove_if(terms.begin(), terms.end(), [](const Term& t) { return t.coef == 0.0 || t.t.op() == CONST; }),
                terms.end());
    if (terms.size() >= 2) {
        const Symbolic f = factorSum(terms, constant);
        if (f.op() != NOOP) return f;
    }
    std::vector<Symbolic> pos, neg;
    for (const Term& t : terms) {
        if (t.coef < 0) neg.push_back(scaled(-t.coef, t.t));
        else pos.push_back(scaled(t.coef, t.t));
    }
    if (constant > 0) pos.push_back(Symbolic(constant));
    else if (constant < 0) neg.push_back(Symbolic(-constant));
    if (!neg.empty()) pos.emplace_back(NEG, neg.size() == 1 ? neg.front() : Symbolic(ADD, neg));
    if (pos.empty()) return Symbolic(0.0);
    return pos.size() == 1 ? pos.front() : Symbolic(ADD, pos);
}

Symbolic normalizeSum(const Symbolic& x) {
    std::vector<Term> terms;
    std::unordered_map<hash_t, size_t> at;
    double constant = 0.0;
    collectTerms(x, 1.0, terms, at, constant);
    return buildSum(std::move(terms), constant);
}

Symbolic normalize(const Symbolic& y) {
    switch (y.op()) {
        case ADD:
        case NEG:
            return normalizeSum(y);
        case MUL:
        case MULC:
        case DIV:
        case SQRT:
            return normalizeProduct(y);
        default:
            return y;
    }
}

}  // namespace

Symbolic foldConstants(const Symbolic& x) {
    return traverseGenerate<Symbolic>(
        x, [](const Symbolic& y) { return y; },
        [](const Symbolic& y, std::vector<Symbolic>& ops) { return y.numChilds() ? fold(y, ops) : y; }, true);
}

Symbolic simplify(const Symbolic& x) {
    return traverseGenerate<Symbolic>(
        x, [](const Symbolic& y) { return y; },
        [](const Symbolic& y, std::vector<Symbolic>& ops) {
            if (!y.numChilds()) return y;
            const Symbolic f = fold(y, ops);
            if (f.op() == CONST || statement(f.op())) return f;
            const Symbolic n = normalize(f);
            return isSmallConstant(n.ahash()) && n.op() != CONST ? Symbolic((double)n.ahash()) : n;
        },
        true);
}

}  // namespace Sym
