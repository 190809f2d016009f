#include "Assembly.hpp"

#include <algorithm>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>

#include "ComputeKernel.hpp"
#include "Parallel.hpp"
#include "Template.hpp"
#include "Timer.hpp"
#include "Traverse.hpp"

namespace Sym {

int HessianAssembly::addInstance(int numVars, int group, int dofs) {
    if (numVars <= 0) throw std::invalid_argument("HessianAssembly: instance without variables");
    if (dofs <= 0 || numVars % dofs) throw std::invalid_argument("HessianAssembly: dofs must divide the variable count");
    for (auto& I : inst_)
        if (I.group == group) throw std::invalid_argument("HessianAssembly: two instances share a variable group");
    Instance I;
    I.numVars = numVars;
    I.group = group;
    I.dofs = dofs;
    I.firstElem = elemTmpl_.size();
    inst_.push_back(std::move(I));
    return (int)inst_.size() - 1;
}

void HessianAssembly::addElement(const Symbolic& energy) {
    if (inst_.empty()) throw std::logic_error("HessianAssembly: addElement before addInstance");
    Instance& I = inst_.back();
    const int group = I.group;
    ElementShape s;
    analyzeShape(energy, [group](const Symbolic& v) { return v.variable()[1] == group; }, s);
    const ExprTemplate& t = hessianTemplateFor(s);
    elemTmpl_.push_back(t.id);
    elemArg_.push_back(args_.size());
    for (auto& v : s.diffLeaves) {
        const int id = v.variable()[0];
        if (id < 0 || id >= I.numVars) throw std::invalid_argument("HessianAssembly: variable outside its instance");
        args_.push_back({0.0, id, 1});
    }
    for (auto& v : s.otherLeaves) {
        if (v.op() == CONST) args_.push_back({v.constant(), 0, 0});
        else throw std::invalid_argument("HessianAssembly: element reads a variable of another group");
    }
    ++I.numElems;
}

int HessianAssembly::defineShape(const Symbolic& energy, int numLocals, int numParams) {
    ElementShape s;
    analyzeShape(energy, [](const Symbolic& v) { return v.variable()[1] == ELEMENT_LOCAL; }, s);
    Shape sh;
    sh.tmpl = hessianTemplateFor(s).id;
    sh.numLocals = numLocals;
    sh.numParams = numParams;
    for (auto& v : s.diffLeaves) {
        const int k = v.variable()[0];
        if (k < 0 || k >= numLocals) throw std::invalid_argument("HessianAssembly: local variable out of range");
        sh.diffLocal.push_back(k);
    }
    for (auto& v : s.otherLeaves) {
        if (v.op() == VAR) {
            const auto id = v.variable();
            if (id[1] != ELEMENT_PARAM || id[0] < 0 || id[0] >= numParams)
                throw std::invalid_argument("HessianAssembly: element shape reads an unknown variable");
            sh.otherLocal.push_back(-1);
            sh.otherParam.push_back(id[0]);
            sh.otherConst.push_back(0.0);
        } else {
            sh.otherLocal.push_back(-1);
            sh.otherParam.push_back(-1);
            sh.otherConst.push_back(v.constant());
        }
    }
    shapes_.push_back(std::move(sh));
    return (int)shapes_.size() - 1;
}

int HessianAssembly::defineMatrixShape(int k, const Symbolic* entries, const int* patternLocals, int numLocals,
                                       int numParams) {
    if (k <= 0 || k > 255 || !entries || !patternLocals || numLocals < k || numParams < 0)
        throw std::invalid_argument("HessianAssembly: bad element matrix shape");
    // template slots: the k row locals, the remaining locals in order, the parameters
    std::vector<int> slotOfLocal((size_t)numLocals, -1);
    Shape sh;
    sh.numLocals = numLocals;
    sh.numParams = numParams;
    for (int i = 0; i < k; ++i) {
        const int q = patternLocals[i];
        if (q < 0 || q >= numLocals || slotOfLocal[(size_t)q] >= 0)
            throw std::invalid_argument("HessianAssembly: element matrix rows must be distinct locals");
        slotOfLocal[(size_t)q] = i;
        sh.diffLocal.push_back(q);
    }
    int slot = k;
    for (int q = 0; q < numLocals; ++q)
        if (slotOfLocal[(size_t)q] < 0) {
            slotOfLocal[(size_t)q] = slot++;
            sh.otherLocal.push_back(q);
            sh.otherParam.push_back(-1);
            sh.otherConst.push_back(0.0);
        }
    for (int j = 0; j < numParams; ++j) {
        sh.otherLocal.push_back(-1);
        sh.otherParam.push_back(j);
        sh.otherConst.push_back(0.0);
    }
    const int firstParamSlot = slot;
    auto toArgs = [&](const Symbolic& x) {
        return traverseGenerate<Symbolic>(x, [&](const Symbolic& y, std::vector<Symbolic>& childs) {
            if (y.op() == VAR) {
                const auto v = y.variable();
                if (v[1] == ELEMENT_LOCAL && v[0] >= 0 && v[0] < numLocals)
                    return Symbolic(slotOfLocal[(size_t)v[0]], TEMPLATE_ARG);
                if (v[1] == ELEMENT_PARAM && v[0] >= 0 && v[0] < numParams)
                    return Symbolic(firstParamSlot + v[0], TEMPLATE_ARG);
                throw std::invalid_argument("HessianAssembly: element matrix reads an unknown variable");
            }
            if (y.numChilds() == 0) return y;
            return Symbolic(y.op(), childs);
        });
    };
    auto t = std::make_unique<ExprTemplate>();
    t->numArgs = firstParamSlot + numParams;
    t->numDiffVars = k;
    t->entryToOutput.assign((size_t)k * k, -1);
    for (int i = 0; i < k; ++i)
        for (int j = i; j < k; ++j) {
            const Symbolic& a = entries[(size_t)i * k + j];
            const Symbolic& b = entries[(size_t)j * k + i];
            if (a.ahash() != b.ahash()) throw std::invalid_argument("HessianAssembly: element matrix is not symmetric");
            t->entryToOutput[(size_t)i * k + j] = t->entryToOutput[(size_t)j * k + i] = t->numOutputs++;
            t->outputs.push_back(toArgs(a));
        }
    // a signature no energy shape has: matrix templates are never shared by lookup
    static int64_t matrixTemplates = 0;
    t->signature = {std::numeric_limits<int64_t>::min(), ++matrixTemplates};
    t->signatureHash = hash(std::vector<hash_t>{(hash_t)0x6d61747269785ull, (hash_t)matrixTemplates});
    sh.tmpl = registerTemplate(std::move(t));
    shapes_.push_back(std::move(sh));
    return (int)shapes_.size() - 1;
}

void HessianAssembly::setNodePattern() {
    if (inst_.empty()) throw std::logic_error("HessianAssembly: setNodePattern before addInstance");
    inst_.back().nodePattern = true;
}

SparseMatrix<double> HessianAssembly::reportedPattern(size_t k, SparseMatrix<double>&& full) const {
    const Instance& I = inst_[k];
    if (!I.nodePattern || I.dofs == 1) return std::move(full);
    const int n = I.numVars / I.dofs;
    std::vector<int> outer(full.outer().begin(), full.outer().begin() + n + 1);
    SparseMatrix<double> P(n, n);
    P.setStructure(std::move(outer), full.releaseInner());
    return P;
}

void HessianAssembly::addElement(int shape, const int* vars, const double* params) {
    if (inst_.empty()) throw std::logic_error("HessianAssembly: addElement before addInstance");
    if (shape < 0 || (size_t)shape >= shapes_.size()) throw std::invalid_argument("HessianAssembly: unknown shape");
    Instance& I = inst_.back();
    const Shape& sh = shapes_[shape];
    elemTmpl_.push_back(sh.tmpl);
    elemArg_.push_back(args_.size());
    for (int k : sh.diffLocal) {
        const int id = vars[k];
        if (id < 0 || id >= I.numVars) throw std::invalid_argument("HessianAssembly: variable outside its instance");
        args_.push_back({0.0, id, 1});
    }
    for (size_t q = 0; q < sh.otherParam.size(); ++q) {
        if (sh.otherLocal[q] >= 0) {
            const int id = vars[sh.otherLocal[q]];
            if (id < 0 || id >= I.numVars) throw std::invalid_argument("HessianAssembly: variable outside its instance");
            args_.push_back({0.0, id, 1});
        } else {
            args_.push_back({sh.otherParam[q] >= 0 ? params[sh.otherParam[q]] : sh.otherConst[q], 0, 0});
        }
    }
    ++I.numElems;
}

void HessianAssembly::addConstant(int row, int col, double value) {
    if (inst_.empty()) throw std::logic_error("HessianAssembly: addConstant before addInstance");
    Instance& I = inst_.back();
    if (row < 0 || col < 0 || row >= I.numVars || col >= I.numVars)
        throw std::invalid_argument("HessianAssembly: constant outside the Hessian");
    I.cRow.push_back(row);
    I.cCol.push_back(col);
    I.cVal.push_back(value);
}

void HessianAssembly::addNodalConstant(int numNodes, const int* nodes, const double* M) {
    if (inst_.empty()) throw std::logic_error("HessianAssembly: addNodalConstant before addInstance");
    Instance& I = inst_.back();
    const int N = I.numVars / I.dofs;
    if (numNodes <= 0) throw std::invalid_argument("HessianAssembly: nodal constant without nodes");
    for (int i = 0; i < numNodes; ++i)
        if (nodes[i] < 0 || nodes[i] >= N) throw std::invalid_argument("HessianAssembly: nodal constant outside the nodes");
    I.ncNodes.insert(I.ncNodes.end(), nodes, nodes + numNodes);
    I.ncStart.push_back(I.ncNodes.size());
    I.ncVal.insert(I.ncVal.end(), M, M + (size_t)numNodes * numNodes);
    I.ncValStart.push_back(I.ncVal.size());
}

namespace {

// Store only base columns; a column equal to a stored one plus a constant reads that one.
void compressColumns(std::vector<std::vector<uint32_t>>& cols, std::vector<ColumnRef>& access,
                     const std::vector<char>& compressible) {
    std::vector<std::vector<uint32_t>> kept;
    std::vector<int> keptSrc;
    access.clear();
    for (size_t c = 0; c < cols.size(); ++c) {
        int found = -1;
        long off = 0;
        if (compressible[c] && !cols[c].empty())
            for (size_t k = 0; k < kept.size() && found < 0; ++k) {
                if (!compressible[keptSrc[k]]) continue;
                const long d = (long)cols[c][0] - (long)kept[k][0];
                bool ok = true;
                for (size_t i = 1; i < cols[c].size() && ok; ++i) ok = (long)cols[c][i] - (long)kept[k][i] == d;
                if (ok && d > std::numeric_limits<int>::min() / 2 && d < std::numeric_limits<int>::max() / 2) {
                    found = (int)k;
                    off = d;
                }
            }
        if (found >= 0) {
            access.push_back({found, (int)off});
        } else {
            access.push_back({(int)kept.size(), 0});
            kept.push_back(std::move(cols[c]));
            keptSrc.push_back((int)c);
        }
    }
    cols.swap(kept);
}

}  // namespace

void HessianAssembly::inputLayout(ProgramInputs& io, std::vector<size_t>& inOff) const {
    io = ProgramInputs();
    for (auto& I : inst_) {
        io.inputSizes.push_back((size_t)I.numVars);
        io.inputOffsets[I.group] = io.inputDataSize;
        io.inputLengths[I.group] = (size_t)I.numVars;
        io.inputDataSize += (size_t)I.numVars;
    }
    if (io.inputDataSize > std::numeric_limits<uint32_t>::max() / 2)
        throw std::length_error("HessianAssembly: too many input values");
    inOff.resize(inst_.size());
    for (size_t k = 0; k < inst_.size(); ++k) inOff[k] = io.inputOffsets.at(inst_[k].group);
}

// Pattern of instance k: every element contributes all numDiffVars^2 (row, col) pairs of its
// variables; constants add (row, col) pairs; duplicates are summed in CSC order (counting sort
// by column, stable sort by row, as setFromTriplets does).
void HessianAssembly::instancePattern(size_t k, InstancePattern& out) const {
    const Instance& I = inst_[k];
    const int nvar = I.numVars;
    std::vector<uint32_t> tr, tc;
    std::vector<uint64_t> src;  // element contribution (e << 16 | slot), or CONST_FLAG | constant index
    const int D = I.dofs, Nn = I.numVars / I.dofs;
    size_t est = I.cRow.size() + (size_t)D * I.ncVal.size();
    for (size_t e = I.firstElem; e < I.firstElem + I.numElems; ++e) {
        const int nd = getTemplate(elemTmpl_[e]).numDiffVars;
        est += (size_t)nd * nd;
    }
    tr.reserve(est);
    tc.reserve(est);
    src.reserve(est);
    const uint64_t CONST_FLAG = uint64_t(1) << 63;
    for (size_t e = I.firstElem; e < I.firstElem + I.numElems; ++e) {
        const ExprTemplate& t = getTemplate(elemTmpl_[e]);
        const int nd = t.numDiffVars;
        const Arg* a = &args_[elemArg_[e]];
        if (I.nodePattern)
            for (int i = 0; i < nd; ++i)
                if (a[i].id >= Nn) throw std::invalid_argument("HessianAssembly: node pattern with a row beyond dof 0");
        for (int i = 0; i < nd; ++i)
            for (int j = 0; j < nd; ++j) {
                const int slot = t.entryToOutput[(size_t)i * nd + j];
                tr.push_back((uint32_t)a[i].id);
                tc.push_back((uint32_t)a[j].id);
                src.push_back((uint64_t)e << 16 | (uint64_t)slot);
            }
    }
    if (I.nodePattern && (!I.ncVal.empty() || std::any_of(I.cRow.begin(), I.cRow.end(), [&](int r) { return r >= Nn; }) ||
                          std::any_of(I.cCol.begin(), I.cCol.end(), [&](int c) { return c >= Nn; })))
        throw std::invalid_argument("HessianAssembly: node pattern with constants beyond dof 0");
    for (size_t c = 0; c < I.cRow.size(); ++c) {
        tr.push_back((uint32_t)I.cRow[c]);
        tc.push_back((uint32_t)I.cCol[c]);
        src.push_back(CONST_FLAG | c);
    }
    const uint64_t NODAL_FLAG = uint64_t(1) << 62;
    for (size_t b = 0; b + 1 < I.ncStart.size(); ++b) {
        const int* nodes = &I.ncNodes[I.ncStart[b]];
        const size_t nn = I.ncStart[b + 1] - I.ncStart[b];
        for (int d = 0; d < D; ++d)
            for (size_t i = 0; i < nn; ++i)
                for (size_t j = 0; j < nn; ++j) {
                    tr.push_back((uint32_t)(nodes[i] + d * Nn));
                    tc.push_back((uint32_t)(nodes[j] + d * Nn));
                    src.push_back(CONST_FLAG | NODAL_FLAG | (uint64_t)(I.ncValStart[b] + i * nn + j));
                }
    }
    const size_t m = tr.size();
    std::vector<uint64_t> cstart((size_t)nvar + 1, 0);
    for (size_t q = 0; q < m; ++q) ++cstart[tc[q] + 1];
    for (int j = 0; j < nvar; ++j) cstart[j + 1] += cstart[j];
    std::vector<uint32_t> order(m);
    {
        std::vector<uint64_t> fill(cstart.begin(), cstart.end() - 1);
        for (size_t q = 0; q < m; ++q) order[fill[tc[q]]++] = (uint32_t)q;
    }
    out = InstancePattern();
    std::vector<int> outer((size_t)nvar + 1, 0), inner;
    inner.reserve(m / 2);
    out.contrib.reserve(m);
    for (int j = 0; j < nvar; ++j) {
        auto b = order.begin() + (long)cstart[j], e = order.begin() + (long)cstart[j + 1];
        std::stable_sort(b, e, [&](uint32_t x, uint32_t y) { return tr[x] < tr[y]; });
        outer[j] = (int)inner.size();
        for (auto it = b; it != e;) {
            auto it2 = it;
            uint32_t first = (uint32_t)out.contrib.size(), count = 0;
            double cval = 0.0;
            bool hasConst = false;
            while (it2 != e && tr[*it2] == tr[*it]) {
                const uint64_t s = src[*it2];
                if (s & CONST_FLAG) {
                    if (s & NODAL_FLAG) cval += I.ncVal[(size_t)(s & ~(CONST_FLAG | NODAL_FLAG))];
                    else cval += I.cVal[(size_t)(s & ~CONST_FLAG)];
                    hasConst = true;
                } else {
                    out.contrib.push_back(s);
                    ++count;
                }
                ++it2;
            }
            inner.push_back((int)tr[*it]);
            out.first.push_back(first);
            out.count.push_back(count);
            out.cval.push_back(cval);
            out.hasConst.push_back(hasConst);
            it = it2;
        }
    }
    outer[nvar] = (int)inner.size();
    out.pattern = SparseMatrix<double>(nvar, nvar);
    out.pattern.setStructure(std::move(outer), std::move(inner));
}

Program HessianAssembly::build(const BuildOptions& opt, ProgramInputs& io, std::vector<SparseMatrix<double>>& patterns) const {
    Program prog;
    PhaseTimer timer;
    std::vector<size_t> inOff;
    inputLayout(io, inOff);
    prog.numInputs = io.inputDataSize;

    // ---- element kernels: one per template, elements in insertion order
    std::map<uint32_t, std::vector<uint32_t>> byTemplate;
    std::vector<uint32_t> elemInst(elemTmpl_.size());
    for (size_t k = 0; k < inst_.size(); ++k)
        for (size_t e = inst_[k].firstElem; e < inst_[k].firstElem + inst_[k].numElems; ++e) elemInst[e] = (uint32_t)k;
    for (size_t e = 0; e < elemTmpl_.size(); ++e) byTemplate[elemTmpl_[e]].push_back((uint32_t)e);

    std::vector<uint32_t> elemKernel(elemTmpl_.size()), elemLocal(elemTmpl_.size());
    for (auto& kv : byTemplate) {
        const ExprTemplate& t = getTemplate(kv.first);
        const auto& elems = kv.second;
        const size_t n = elems.size();
        const int nv = t.numDiffVars;
        ProgramKernel P;
        P.n = n;
        P.wave = 0;
        P.userOutputs = false;
        P.numOutSlots = (size_t)t.numOutputs;
        P.interBase = prog.numIntermediates;
        prog.numIntermediates += P.numOutSlots * n;

        // argument slots -> variable columns, constant columns or literals
        std::vector<std::vector<uint32_t>> vcols;
        std::vector<char> compressible;
        std::vector<std::vector<double>> ccols;
        std::vector<Symbolic> argRef(t.numArgs);
        for (int a = 0; a < t.numArgs; ++a) {
            const Arg& first = args_[elemArg_[elems[0]] + a];
            if (first.kind == 1) {
                std::vector<uint32_t> col(n);
                for (size_t i = 0; i < n; ++i) {
                    const Arg& g = args_[elemArg_[elems[i]] + a];
                    if (g.kind != 1) throw std::logic_error("HessianAssembly: argument kinds differ within a template");
                    col[i] = (uint32_t)(inOff[elemInst[elems[i]]] + (size_t)g.id);
                }
                argRef[a] = Symbolic((int)vcols.size(), -3);
                vcols.push_back(std::move(col));
                compressible.push_back(1);
            } else {
                std::vector<double> col(n);
                bool same = true;
                for (size_t i = 0; i < n; ++i) {
                    const Arg& g = args_[elemArg_[elems[i]] + a];
                    if (g.kind != 0) throw std::logic_error("HessianAssembly: argument kinds differ within a template");
                    col[i] = g.value;
                    same = same && col[i] == col[0];
                }
                if (same) {
                    argRef[a] = Symbolic(col[0]);
                } else {
                    argRef[a] = Symbolic((int)ccols.size(), -4);
                    ccols.push_back(std::move(col));
                }
            }
        }
        (void)nv;
        std::vector<Symbolic> assigns;
        assigns.reserve(t.numOutputs);
        for (int k = 0; k < t.numOutputs; ++k)
            assigns.emplace_back(ASSIGN, Symbolic(k, -2), substituteArguments(t.outputs[k], argRef));
        P.code = optimizeTemplate(Symbolic(BLOCK, assigns), opt.simplifyTemplates);

        compressColumns(vcols, P.access, compressible);
        P.numVarSlots = vcols.size();
        P.indexOffset = prog.index.size();
        for (auto& c : vcols) prog.index.insert(prog.index.end(), c.begin(), c.end());
        P.numConstSlots = ccols.size();
        P.constOffset = prog.consts.size();
        for (auto& c : ccols) prog.consts.insert(prog.consts.end(), c.begin(), c.end());

        const uint32_t kid = (uint32_t)prog.kernels.size();
        for (size_t i = 0; i < n; ++i) {
            elemKernel[elems[i]] = kid;
            elemLocal[elems[i]] = (uint32_t)i;
        }
        prog.kernels.push_back(std::move(P));
    }

    timer("assembly: element kernels");
    // ---- per instance: pattern and contribution lists
    struct OutRec {
        uint32_t first, count;  // into contrib
        double cval;
        bool hasConst;
        uint32_t pos;           // result position
    };
    std::vector<InstancePattern> results(inst_.size());
    parallelFor(inst_.size(), [&](size_t k) { instancePattern(k, results[k]); });

    // merge in instance order; contributions become intermediate positions
    std::vector<OutRec> outs;
    std::vector<uint32_t> contrib;
    size_t resultBase = 0, totalOuts = 0, totalContrib = 0;
    for (auto& r : results) {
        totalOuts += r.first.size();
        totalContrib += r.contrib.size();
    }
    outs.reserve(totalOuts);
    contrib.reserve(totalContrib);
    patterns.clear();
    for (size_t k = 0; k < inst_.size(); ++k) {
        InstancePattern& r = results[k];
        const size_t nout = r.pattern.nonZeros();
        if (resultBase + nout > std::numeric_limits<uint32_t>::max())
            throw std::length_error("HessianAssembly: too many outputs");
        const uint32_t cbase = (uint32_t)contrib.size();
        for (size_t q = 0; q < nout; ++q)
            outs.push_back({r.first[q] + cbase, r.count[q], r.cval[q], (bool)r.hasConst[q], (uint32_t)(resultBase + q)});
        for (uint64_t s : r.contrib) {
            const size_t e = (size_t)(s >> 16), slot = (size_t)(s & 0xffff);
            const ProgramKernel& P = prog.kernels[elemKernel[e]];
            contrib.push_back((uint32_t)(prog.numInputs + P.interBase + slot * P.n + elemLocal[e]));
        }
        prog.outStart.push_back(resultBase);
        prog.outLength.push_back(nout);
        io.outputSizes.push_back(nout);
        resultBase += nout;
        patterns.push_back(reportedPattern(k, std::move(r.pattern)));
        r = InstancePattern();
    }

    timer("assembly: patterns");
    // ---- output kernels: one per (number of element contributions, constant term)
    std::map<std::pair<uint32_t, bool>, std::vector<uint32_t>> classes;
    for (size_t o = 0; o < outs.size(); ++o) classes[{outs[o].count, outs[o].hasConst}].push_back((uint32_t)o);
    for (auto& kv : classes) {
        const uint32_t kc = kv.first.first;
        const bool hc = kv.first.second;
        const auto& list = kv.second;
        const size_t n = list.size();
        ProgramKernel P;
        P.n = n;
        P.wave = 1;
        P.userOutputs = true;
        P.numOutSlots = 1;
        std::vector<Symbolic> terms;
        for (uint32_t q = 0; q < kc; ++q) terms.emplace_back((int)q, -3);
        if (hc) terms.emplace_back(0, -4);
        if (terms.empty()) throw std::logic_error("HessianAssembly: output without contributions");
        Symbolic sum = terms.size() == 1 ? terms[0] : Symbolic(ADD, terms);
        P.code = Symbolic(BLOCK, std::vector<Symbolic>{Symbolic(ASSIGN, Symbolic(0, -2), sum)});
        P.numVarSlots = kc;
        P.indexOffset = prog.index.size();
        for (uint32_t q = 0; q < kc; ++q)
            for (size_t i = 0; i < n; ++i) prog.index.push_back(contrib[outs[list[i]].first + q]);
        for (uint32_t q = 0; q < kc; ++q) P.access.push_back({(int)q, 0});
        P.numConstSlots = hc ? 1 : 0;
        P.constOffset = prog.consts.size();
        if (hc)
            for (size_t i = 0; i < n; ++i) prog.consts.push_back(outs[list[i]].cval);
        P.outposOffset = prog.outpos.size();
        for (size_t i = 0; i < n; ++i) prog.outpos.push_back(outs[list[i]].pos);
        prog.kernels.push_back(std::move(P));
    }
    prog.numWaves = prog.kernels.empty() ? 0 : 2;
    timer("assembly: output kernels");
    if (prog.numInputs + prog.numIntermediates > std::numeric_limits<uint32_t>::max())
        throw std::length_error("HessianAssembly: value buffer exceeds 32-bit positions");
    return prog;
}

}  // namespace Sym
