#include "Simplify.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <set>

#include "Traverse.hpp"
#include "Utilities.hpp"

namespace Sym {

namespace {

bool isMul(const Symbolic& x) { return x.op() == MUL || x.op() == MULC; }

Symbolic makePower(const Symbolic& x, const double d) {
    if (x.op() == CONST) return Symbolic(std::pow(x.constant(), d));
    if (x.op() == POW && x[1].op() == CONST) return Symbolic(POW, x[0], Symbolic(d * x[1].constant()));
    return Symbolic(POW, x, Symbolic(d));
}

Symbolic pullConstant(const Symbolic& x_) {
    Symbolic x = flattenMultiplications(x_);
    if (x.op() == MULC) return x;
    if (x.op() != MUL) return Symbolic(MULC, Symbolic(1.), x);
    double f = 1.;
    std::vector<Symbolic> childs;
    for (auto& c : x) {
        if (c.op() == CONST) f *= c.constant();
        else childs.push_back(c);
    }
    return Symbolic(MULC, Symbolic(f), makeSymbolic(MUL, childs));
}

Symbolic simplifyExpression(const Symbolic& expr) {
    if (isSmallConstant(expr.ahash())) return Symbolic((double)expr.ahash());
    if (expr.numChilds() == 0) return expr;
    Symbolic x = optimizeSum(expr);
    x = pullFactor(x);
    x = optimizeProducts(x);
    return x;
}

}  // namespace

Symbolic makeMul(const std::vector<Symbolic>& childs) {
    if (childs.empty()) return Symbolic(1.);
    std::vector<Symbolic> childs2;
    double f = 1.;
    for (auto& c : childs) {
        if (c.op() == CONST) f *= c.constant();
        else if (isSmallConstant(c.ahash())) f *= (double)c.ahash();
        else childs2.push_back(c);
    }
    if (f == 0.0) return Symbolic(0.0);
    if (f != 1.) childs2.push_back(Symbolic(f));
    if (childs2.empty()) return Symbolic(1.);
    if (childs2.size() == 1) return childs2.front();
    return Symbolic(MUL, childs2);
}

Symbolic makeAddition(std::vector<Symbolic> operands) {
    double constant = 0.0;
    operands.erase(std::remove_if(operands.begin(), operands.end(),
                                  [&](const Symbolic& x) {
                                      if (x.op() == CONST) {
                                          constant += x.constant();
                                          return true;
                                      }
                                      return false;
                                  }),
                   operands.end());
    if (constant != 0) operands.push_back(Symbolic(constant));
    if (operands.empty()) return Symbolic(0.0);
    if (operands.size() == 1) return operands.front();
    return Symbolic(ADD, operands);
}

Symbolic makeSymbolic(const unsigned char op, const std::vector<Symbolic>& operands) {
    Symbolic ret;
    switch (op) {
        case ADD:
            ret = makeAddition(operands);
            break;
        case MUL:
        case MULC:
            ret = makeMul(operands);
            break;
        default:
            ret = Symbolic(op, operands);
    }
    if (op != ASSIGN && op != DEFINE && op != BLOCK && isSmallConstant(ret.ahash())) return Symbolic((double)ret.ahash());
    return ret;
}

Symbolic flattenAdditions(const Symbolic& expr) {
    if (expr.op() != ADD) return expr;
    std::vector<Symbolic> operands;
    preOrderTraverse(expr, [&](const Symbolic& x) {
        if (!x.ahash()) return false;
        if (x.op() != ADD) {
            operands.push_back(x);
            return false;
        }
        return true;
    });
    Symbolic ret(ADD, operands);
    if (ret.ahash() == 0) ret = Symbolic(0.0);
    else if (ret.numChilds() == 1) ret = Symbolic(ret[0]);
    return ret;
}

Symbolic flattenMultiplications(const Symbolic& expr) {
    if (!isMul(expr)) return expr;
    std::vector<Symbolic> operands;
    preOrderTraverse(expr, [&](const Symbolic& x) {
        if (isSmallConstant(x.ahash())) {
            if (x.ahash() != 1) operands.push_back(Symbolic((double)x.ahash()));
            return false;
        }
        if (!isMul(x)) {
            operands.push_back(x);
            return false;
        }
        return true;
    });
    return makeSymbolic(MUL, operands);
}

Symbolic flattenDivAndMul(const Symbolic& expr) {
    std::vector<Symbolic> childs;
    switch (expr.op()) {
        case MUL:
        case MULC:
            for (auto& c0 : expr) {
                Symbolic c = flattenDivAndMul(c0);
                if (isMul(c))
                    for (auto& cc : c) childs.push_back(cc);
                else childs.push_back(c);
            }
            break;
        case SQRT: {
            Symbolic rad = flattenDivAndMul(expr[0]);
            if (isMul(rad))
                for (auto& c : rad) childs.push_back(makePower(c, 0.5));
            else childs.push_back(makePower(rad, 0.5));
            break;
        }
        case DIV: {
            Symbolic nom = flattenDivAndMul(expr[0]);
            Symbolic den = flattenDivAndMul(expr[1]);
            if (isMul(nom))
                for (auto& c : nom) childs.push_back(c);
            else childs.push_back(nom);
            if (isMul(den))
                for (auto& c : den) childs.push_back(makePower(c, -1.));
            else childs.push_back(makePower(den, -1.));
            break;
        }
        default:
            return expr;
    }
    return makeMul(childs);
}

Symbolic optimizeProducts(const Symbolic& x) {
    Symbolic xf = flattenDivAndMul(x);
    if (!isMul(xf)) return x;

    std::map<Symbolic, double, AlgebraicHashFunctor> factors;
    for (auto& c : xf) {
        if (c.op() == POW) {
            if (c[1].op() != CONST) return x;
            factors[c[0]] += c[1].constant();
        } else factors[c] += 1.;
    }

    std::vector<Symbolic> nom, den, nomR, denR;
    double f = 1.;
    for (auto& c : factors) {
        if (c.first.op() == CONST) {
            f *= std::pow(c.first.constant(), c.second);
        } else if (c.second < 0) {
            const double e = -c.second;
            if (std::floor(e) == e) {
                for (int i = 0; i < (int)e; ++i) den.push_back(c.first);
            } else if (std::floor(2 * e) == 2 * e) {
                for (int i = 0; i < (int)std::floor(e); ++i) den.push_back(c.first);
                denR.push_back(c.first);
            } else return x;
        } else if (c.second > 0) {
            const double e = c.second;
            if (std::floor(e) == e) {
                for (int i = 0; i < (int)e; ++i) nom.push_back(c.first);
            } else if (std::floor(2 * e) == 2 * e) {
                for (int i = 0; i < (int)std::floor(e); ++i) nom.push_back(c.first);
                nomR.push_back(c.first);
            } else return x;
        }
    }

    if (!nomR.empty()) {
        if (denR.empty()) nom.emplace_back(SQRT, makeMul(nomR));
        else nom.emplace_back(SQRT, makeMul(nomR) / makeMul(denR));
    } else if (!denR.empty()) {
        den.emplace_back(SQRT, makeMul(denR));
    }

    if (f != 1.) nom.push_back(Symbolic(f));
    if (nom.empty()) nom.push_back(Symbolic(1.));
    Symbolic ret = makeMul(nom);
    if (!ret.ahash()) return Symbolic(0.0);
    if (!den.empty()) ret = ret / makeMul(den);
    return ret;
}

Symbolic optimizeSum(const Symbolic& x0) {
    Symbolic x = flattenAdditions(x0);
    if (x.op() != ADD) return x;

    std::map<Symbolic, double, AlgebraicHashFunctor> multiples;
    for (auto& c0 : x) {
        Symbolic c = pullConstant(c0);
        multiples[c[1]] += c[0].constant();
    }
    std::vector<Symbolic> childs;
    for (auto& c : multiples) {
        if (c.second != 0.0) {
            if (c.second == 1.) childs.push_back(c.first);
            else childs.push_back(flattenMultiplications(Symbolic(c.second) * c.first));
        }
    }
    return makeSymbolic(ADD, childs);
}

Symbolic pullFactor(Symbolic x) {
    if (x.op() != ADD) return x;
    if (x.numChilds() == 1) return x[0];
    if (x.numChilds() == 0) return Symbolic(0.0);
    if (isSmallConstant(x.algebraicHash())) return Symbolic((double)x.algebraicHash());

    std::vector<std::vector<Symbolic>> summands;
    summands.reserve(x.numChilds());
    for (const auto& c : x) {
        if (isMul(c)) {
            Symbolic fc = flattenMultiplications(c);
            if (fc.op() == MUL) summands.emplace_back(fc.begin(), fc.end());
            else summands.push_back({fc});
        } else summands.push_back({c});
    }

    // the factor shared by the most summands (weighted by its complexity)
    std::map<Symbolic, std::set<int>, AlgebraicHashFunctor> factors;
    for (int i = 0; i < (int)summands.size(); ++i)
        for (auto& y : summands[i]) factors[y].insert(i);

    auto maxIt = std::max_element(factors.begin(), factors.end(), [](const auto& a, const auto& b) {
        const int scoreA = std::max((int)a.second.size() - 2, 0) * (a.first.complexity() + 1);
        const int scoreB = std::max((int)b.second.size() - 2, 0) * (b.first.complexity() + 1);
        return scoreA < scoreB;
    });
    if (maxIt->second.size() < 2) return x;

    auto it = maxIt->second.begin();
    std::vector<Symbolic> factored, rest;
    for (int i = 0; i < (int)summands.size(); ++i) {
        if (it != maxIt->second.end() && i == *it) {
            auto its = std::find(summands[i].begin(), summands[i].end(), maxIt->first);
            assert(its != summands[i].end());
            summands[i].erase(its);
            if (summands[i].empty()) factored.push_back(Symbolic(1.));
            else factored.emplace_back(MUL, summands[i]);
            ++it;
        } else {
            rest.emplace_back(MUL, summands[i]);
        }
    }
    rest.push_back(maxIt->first * pullFactor(makeSymbolic(ADD, factored)));
    return pullFactor(flattenAdditions(makeSymbolic(ADD, rest)));
}

Symbolic findNegative(const Symbolic& expr) {
    if (expr.op() == MULC && expr[0].op() == CONST && expr[0].constant() == -1) return Symbolic(NEG, expr[1]);
    if (expr.op() == MUL) {
        std::vector<Symbolic> childs;
        int cnt = 0;
        for (auto& c : expr) {
            if (c.op() == CONST && c.constant() == -1.) ++cnt;
            else childs.push_back(c);
        }
        if (cnt % 2) return Symbolic(NEG, makeMul(childs));
        return makeMul(childs);
    } else if (expr.op() == ADD) {
        std::vector<Symbolic> positive, negative;
        for (auto& c : expr) {
            if (c.op() == NEG) negative.push_back(c[0]);
            else if (c.op() == MULC && c[0].op() == CONST && c[0].constant() == -1.) negative.push_back(c[1]);
            else positive.push_back(c);
        }
        if (!negative.empty()) positive.emplace_back(NEG, makeAddition(negative));
        return makeAddition(positive);
    }
    return expr;
}

Symbolic simplify(const Symbolic& expr) {
    return traverseGenerate<Symbolic>(
        expr, [](const Symbolic& x) { return simplifyExpression(x); },
        [](const Symbolic& x, std::vector<Symbolic>& childs) {
            if (x.numChilds() == 0) return x;
            return findNegative(simplifyExpression(makeSymbolic(x.op(), childs)));
        },
        true);
}

Symbolic removeConstantExpressions(const Symbolic& x) {
    return traverseGenerate<Symbolic>(
        x, [](const Symbolic& y) { return y; },
        [](const Symbolic& y, std::vector<Symbolic>& childs) {
            if (y.numChilds() == 0) return y;
            return makeSymbolic(y.op(), childs);
        },
        true);
}

Symbolic referenceRedundant(const Symbolic& x) { return removeConstantExpressions(x); }

}  // namespace Sym
