#include "Template.hpp"

#include <stdexcept>
#include <unordered_map>

#include "Traverse.hpp"

namespace Sym {

namespace {

struct Registry {
    std::vector<std::unique_ptr<ExprTemplate>> all;
    std::unordered_multimap<hash_t, uint32_t> bySignature;
};

Registry& registry() {
    static Registry* r = new Registry();  // lives for the whole process
    return *r;
}

}  // namespace

uint32_t registerTemplate(std::unique_ptr<ExprTemplate> t) {
    auto& r = registry();
    t->id = (uint32_t)r.all.size();
    r.bySignature.emplace(t->signatureHash, t->id);
    r.all.push_back(std::move(t));
    return r.all.back()->id;
}

const ExprTemplate& getTemplate(uint32_t id) {
    auto& r = registry();
    if (id >= r.all.size()) throw std::out_of_range("unknown expression template id");
    return *r.all[id];
}

const ExprTemplate* findTemplate(hash_t signatureHash, const std::vector<int64_t>& signature) {
    auto& r = registry();
    auto range = r.bySignature.equal_range(signatureHash);
    for (auto it = range.first; it != range.second; ++it) {
        const ExprTemplate& t = *r.all[it->second];
        if (t.signature == signature) return &t;
    }
    return nullptr;
}

Symbolic substituteArguments(const Symbolic& x, const std::vector<Symbolic>& args) {
    return traverseGenerate<Symbolic>(x, [&](const Symbolic& y, std::vector<Symbolic>& childs) {
        if (y.op() == VAR) {
            auto v = y.variable();
            if (v[1] == TEMPLATE_ARG) return args.at((size_t)v[0]);
            return y;
        }
        if (y.numChilds() == 0) return y;
        bool same = true;
        for (unsigned i = 0; i < y.numChilds(); ++i)
            if (childs[i].id() != y[i].id()) {
                same = false;
                break;
            }
        if (same) return y;
        if (y.op() == INSTANCE) return Symbolic::instance(y.tag(), childs, 0);
        return Symbolic(y.op(), childs);
    });
}

Symbolic expandInstance(const Symbolic& inst) {
    const ExprTemplate& t = getTemplate(inst.tag());
    if ((int)inst.numChilds() != t.numOutputs + t.numArgs)
        throw std::logic_error("template instance has the wrong number of children");
    std::vector<Symbolic> args(inst.begin() + t.numOutputs, inst.end());
    std::vector<Symbolic> assigns;
    assigns.reserve(t.numOutputs);
    // one substitution map for all outputs keeps shared sub-expressions shared
    std::unordered_map<long long, Symbolic, IdentityHash<long long>> memo;
    for (int k = 0; k < t.numOutputs; ++k) {
        Symbolic e = traverseGenerateMap<Symbolic>(
            t.outputs[k],
            [&](const Symbolic& y, std::vector<Symbolic>& childs) {
                if (y.op() == VAR) {
                    auto v = y.variable();
                    if (v[1] == TEMPLATE_ARG) return args.at((size_t)v[0]);
                    return y;
                }
                if (y.numChilds() == 0) return y;
                return Symbolic(y.op(), childs);
            },
            memo, false);
        assigns.emplace_back(ASSIGN, inst[k], e);
    }
    return Symbolic(BLOCK, assigns);
}

}  // namespace Sym
