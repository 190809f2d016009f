// Fused assembly form: tiling of owner nodes, tile pool layout and validation.
#include "Fused.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <stdexcept>

#include "Assembly.hpp"
#include "ComputeKernel.hpp"
#include "Parallel.hpp"
#include "Template.hpp"
#include "Timer.hpp"

namespace Sym {

namespace {

constexpr uint32_t NONE = 0xffffffffu;
constexpr int HEADER = FUSED_HEADER;  // fixed header words before the per-template triples

uint32_t u32(size_t v, const char* what) {
    if (v > std::numeric_limits<uint32_t>::max()) throw std::length_error(std::string("fused form: ") + what + " exceeds 32 bits");
    return (uint32_t)v;
}

// Per-template layout shared by every element of the template.
struct TemplateLayout {
    uint32_t tmplId = 0;
    std::vector<int> argGather;   // template argument -> gathered index, or -1
    std::vector<int> argParam;    // template argument -> parameter column, or -1
    std::vector<double> literal;  // uniform constant of a constant argument
    std::vector<char> uniform;
    std::vector<int> gatherNode, gatherDof;
    int numNodes = 0, numParams = 0, numOutputs = 0;
    // block sums: the differentiated nodes (layout node of each, first-use order), the value
    // slot of node pair (p, q), p <= q, at blockBase[p * diffNode.size() + q], and the template
    // output held by each value slot
    std::vector<int> diffNode, blockBase, slotOutput;
    std::vector<int> diffArg;  // block sums: an argument of each differentiated node
    int valueStride() const { return numOutputs | 1; }
    // LDS bytes of one element besides its nodes' coordinates: parameters, values, slots
    size_t elementBytes() const { return 8 * (size_t)(numParams + valueStride()) + 2 * (size_t)numNodes; }
};

// Tiles of one instance, pool offsets relative to the instance's chunk.
struct InstanceTiles {
    std::vector<uint32_t> pool;
    std::vector<double> consts;
    std::vector<size_t> header;  // tile header offsets in pool
    std::vector<double> cost;    // per tile: outputs + element values
    std::vector<size_t> values;  // per tile: element values (its image in the split form)
    size_t evaluations = 0, contributions = 0, lds = 0, elementLds = 0, stagedLds = 0, outputs = 0;
    int maxPairs = 0;  // the most slot pairs of an output
    size_t maxBlocks = 0;
};

// The Hessian of an instance is not made of node blocks (or exceeds the block records'
// fields): the program is built with one contribution list per output instead.
struct NotBlockStructured : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// Block value layout of a template over the differentiated arguments 0..nd-1 of `t` for D
// dofs per node; false when the arguments are not every dof 0..D-1 of their nodes, each once.
bool blockLayout(TemplateLayout& tl, const ExprTemplate& t, int D) {
    const int nd = t.numDiffVars;
    if (nd <= 0 || (int)t.entryToOutput.size() != nd * nd) return false;
    std::vector<int> nodeIdx(nd), dofOf(nd);
    tl.diffNode.clear();
    tl.diffArg.clear();
    for (int i = 0; i < nd; ++i) {
        const int g = i < (int)tl.argGather.size() ? tl.argGather[i] : -1;
        if (g < 0) return false;
        const int j = tl.gatherNode[g];
        int p = (int)(std::find(tl.diffNode.begin(), tl.diffNode.end(), j) - tl.diffNode.begin());
        if (p == (int)tl.diffNode.size()) {
            tl.diffNode.push_back(j);
            tl.diffArg.push_back(i);
        }
        nodeIdx[i] = p;
        dofOf[i] = tl.gatherDof[g];
    }
    const int n = (int)tl.diffNode.size();
    if (n * D != nd) return false;
    std::vector<int> argOf(nd, -1);  // (p, d) -> argument
    for (int i = 0; i < nd; ++i) {
        if (dofOf[i] < 0 || dofOf[i] >= D || argOf[nodeIdx[i] * D + dofOf[i]] >= 0) return false;
        argOf[nodeIdx[i] * D + dofOf[i]] = i;
    }
    tl.blockBase.assign((size_t)n * n, -1);
    tl.slotOutput.clear();
    for (int p = 0; p < n; ++p)
        for (int q = p; q < n; ++q) {
            tl.blockBase[(size_t)p * n + q] = (int)tl.slotOutput.size();
            for (int da = 0; da < D; ++da)
                for (int db = p == q ? da : 0; db < D; ++db) {
                    const int i = argOf[p * D + da], j = argOf[q * D + db];
                    const int o = t.entryToOutput[(size_t)i * nd + j];
                    if (o < 0) return false;
                    tl.slotOutput.push_back(o);
                }
        }
    return true;
}

}  // namespace

std::string FusedProgram::summary() const {
    char buf[256];
    std::snprintf(buf, sizeof(buf),
                  "fused tiles=%zu lds=%zu element_lds=%zu recompute=%.3f templates=%zu contributions=%zu pool=%zu",
                  numTiles, ldsBytes, elementLdsBytes, recompute(), templates.size(), numContributions, pool.size());
    return buf;
}

FusedProgram HessianAssembly::buildFused(const BuildOptions& opt, const FusedOptions& fo, ProgramInputs& io,
                                         std::vector<SparseMatrix<double>>& patterns) const {
    if (fo.blockSums) {
        try {
            return buildFusedForm(opt, fo, io, patterns, true);
        } catch (const NotBlockStructured&) {
        }
    }
    return buildFusedForm(opt, fo, io, patterns, false);
}

FusedProgram HessianAssembly::buildFusedForm(const BuildOptions& opt, const FusedOptions& fo, ProgramInputs& io,
                                             std::vector<SparseMatrix<double>>& patterns, bool blocks) const {
    PhaseTimer timer;
    const int T = fo.threadsPerBlock;
    if (T <= 0 || T > 1024 || T % 64) throw std::invalid_argument("fused form: threads per block must be a multiple of 64 up to 1024");
    if (fo.ldsBytes > FUSED_MAX_LDS_BYTES) throw std::invalid_argument("fused form: more LDS per workgroup than a CU has (160 KiB)");
    FusedProgram fp;
    std::vector<size_t> inOff;
    inputLayout(io, inOff);
    fp.numInputs = io.inputDataSize;
    fp.numElements = elemTmpl_.size();

    // ---- templates in first-use order, argument kinds, node layouts
    std::map<uint32_t, int> tIndex;
    std::vector<TemplateLayout> L;
    std::vector<int> elemT(elemTmpl_.size());
    std::vector<uint32_t> elemInst(elemTmpl_.size());
    for (size_t k = 0; k < inst_.size(); ++k)
        for (size_t e = inst_[k].firstElem; e < inst_[k].firstElem + inst_[k].numElems; ++e) elemInst[e] = (uint32_t)k;
    std::vector<int> nodeOf;  // scratch: local node index -> node id of the current element
    for (size_t e = 0; e < elemTmpl_.size(); ++e) {
        auto it = tIndex.find(elemTmpl_[e]);
        const ExprTemplate& t = getTemplate(elemTmpl_[e]);
        const Instance& I = inst_[elemInst[e]];
        const int N = I.numVars / I.dofs;
        const Arg* a = &args_[elemArg_[e]];
        if (it == tIndex.end()) {
            it = tIndex.emplace(elemTmpl_[e], (int)L.size()).first;
            TemplateLayout tl;
            tl.tmplId = elemTmpl_[e];
            tl.numOutputs = t.numOutputs;
            std::vector<int> nodes;
            for (int q = 0; q < t.numArgs; ++q) {
                if (a[q].kind == 1) {
                    const int node = a[q].id % N, dof = a[q].id / N;
                    int j = (int)(std::find(nodes.begin(), nodes.end(), node) - nodes.begin());
                    if (j == (int)nodes.size()) nodes.push_back(node);
                    tl.argGather.push_back((int)tl.gatherNode.size());
                    tl.argParam.push_back(-1);
                    tl.gatherNode.push_back(j);
                    tl.gatherDof.push_back(dof);
                    tl.literal.push_back(0.0);
                    tl.uniform.push_back(0);
                } else {
                    tl.argGather.push_back(-1);
                    tl.argParam.push_back(-1);
                    tl.literal.push_back(a[q].value);
                    tl.uniform.push_back(1);
                }
            }
            tl.numNodes = (int)nodes.size();
            L.push_back(std::move(tl));
        }
        TemplateLayout& tl = L[it->second];
        elemT[e] = it->second;
        nodeOf.assign((size_t)tl.numNodes, -1);
        for (int q = 0; q < t.numArgs; ++q) {
            const bool isVar = a[q].kind == 1;
            if (isVar != (tl.argGather[q] >= 0))
                throw std::invalid_argument("fused form: argument kinds differ within an element template");
            if (isVar) {
                const int g = tl.argGather[q];
                const int node = a[q].id % N, dof = a[q].id / N;
                int& slot = nodeOf[(size_t)tl.gatherNode[g]];
                if (slot < 0) slot = node;
                if (slot != node || dof != tl.gatherDof[g])
                    throw std::invalid_argument("fused form: elements of one template have different node layouts");
            } else if (tl.uniform[q] && a[q].value != tl.literal[q]) {
                tl.uniform[q] = 0;
            }
        }
    }
    // ---- block sums: every instance's Hessian made of D x D node blocks (D = dofs, 1 for a
    // node pattern): no constants besides nodal blocks, every template's differentiated
    // arguments all dofs of its nodes, the nodes of an element distinct
    int Db = 0;
    if (blocks) {
        for (auto& I : inst_) {
            const int d = I.nodePattern ? 1 : I.dofs;
            if (!I.cRow.empty() || (Db && d != Db)) throw NotBlockStructured("constants or dofs");
            Db = d;
        }
        // one dof per node: a block is one output, so per-output lists are as short and the
        // block form's pass through LDS only costs
        if (Db < 2 || Db > 3) throw NotBlockStructured("dofs per node");
        for (auto& tl : L)
            if (!blockLayout(tl, getTemplate(tl.tmplId), Db)) throw NotBlockStructured("element template");
        for (size_t e = 0; e < elemTmpl_.size(); ++e) {
            const TemplateLayout& tl = L[(size_t)elemT[e]];
            const Instance& I = inst_[elemInst[e]];
            const int N = I.numVars / I.dofs;
            const Arg* a = &args_[elemArg_[e]];
            const size_t n = tl.diffNode.size();
            for (size_t p = 0; p < n; ++p)
                for (size_t q = 0; q < p; ++q)
                    if (a[tl.diffArg[p]].id % N == a[tl.diffArg[q]].id % N) throw NotBlockStructured("repeated element node");
        }
        for (auto& tl : L) tl.numOutputs = (int)tl.slotOutput.size();
    }
    for (auto& tl : L) {
        const ExprTemplate& t = getTemplate(tl.tmplId);
        std::vector<Symbolic> argRef((size_t)t.numArgs);
        for (int q = 0; q < t.numArgs; ++q) {
            if (tl.argGather[q] >= 0) argRef[q] = Symbolic(tl.argGather[q], -3);
            else if (tl.uniform[q]) argRef[q] = Symbolic(tl.literal[q]);
            else {
                tl.argParam[q] = tl.numParams++;
                argRef[q] = Symbolic(tl.argParam[q], -4);
            }
        }
        std::vector<Symbolic> assigns;
        if (blocks)
            for (size_t s = 0; s < tl.slotOutput.size(); ++s)
                assigns.emplace_back(ASSIGN, Symbolic((int)s, -2), substituteArguments(t.outputs[(size_t)tl.slotOutput[s]], argRef));
        else
            for (int s = 0; s < t.numOutputs; ++s)
                assigns.emplace_back(ASSIGN, Symbolic(s, -2), substituteArguments(t.outputs[s], argRef));
        FusedTemplate ft;
        ft.code = optimizeTemplate(Symbolic(BLOCK, assigns), opt.simplifyTemplates);
        ft.numNodes = tl.numNodes;
        ft.argNode = tl.gatherNode;
        ft.argDof = tl.gatherDof;
        ft.numParams = tl.numParams;
        ft.numOutputs = tl.numOutputs;
        ft.valueStride = tl.valueStride();
        for (int d : tl.gatherDof) fp.stagedDofs = std::max(fp.stagedDofs, d + 1);
        fp.templates.push_back(std::move(ft));
    }
    if (blocks) {
        fp.sums = FusedSums::Blocks;
        fp.blockDofs = Db;
    }
    const size_t NT = L.size();
    const size_t nodeLds = 8 * (size_t)fp.stagedDofs;  // LDS bytes per staged tile node
    for (auto& tl : L)
        if (tl.elementBytes() + nodeLds * (size_t)tl.numNodes > fo.ldsBytes)
            throw std::invalid_argument("fused form: one element needs more LDS than a tile holds");
    timer("fused: templates");

    // ---- per instance: pattern and tiles
    std::vector<InstanceTiles> tiles(inst_.size());
    std::vector<SparseMatrix<double>> pats(inst_.size());
    // Tiles grow until the workgroup's whole LDS (element values, staged arguments, recipe and
    // output constants) would exceed fo.ldsBytes; recipe and constants are charged per owned
    // node by an upper bound of what it adds (its count bytes, u16 slots, constants, two run
    // words per dof column and its share of the 64-output chunk records), so the tile built
    // never needs more than the budget.
    // margin: run terminator, chunk and word rounding, slot table padding per template
    const double ldsBudget = (double)fo.ldsBytes - 64.0 - 16.0 * (double)L.size();
    parallelFor(inst_.size(), [&](size_t k) {
        const Instance& I = inst_[k];
        InstancePattern P;
        instancePattern(k, P);
        const int D = I.dofs, N = I.numVars / I.dofs;
        const std::vector<int>& colptr = P.pattern.outer();
        const size_t E = I.numElems, e0 = I.firstElem;

        // element nodes (by local node index) and node -> incident elements
        std::vector<uint32_t> enStart(E + 1, 0), en;
        for (size_t i = 0; i < E; ++i) enStart[i + 1] = enStart[i] + (uint32_t)L[elemT[e0 + i]].numNodes;
        en.assign(enStart[E], 0);
        for (size_t i = 0; i < E; ++i) {
            const TemplateLayout& tl = L[elemT[e0 + i]];
            const Arg* a = &args_[elemArg_[e0 + i]];
            for (size_t q = 0; q < tl.argGather.size(); ++q)
                if (tl.argGather[q] >= 0) en[enStart[i] + tl.gatherNode[tl.argGather[q]]] = (uint32_t)(a[q].id % N);
        }
        std::vector<uint32_t> incStart((size_t)N + 1, 0), inc;
        for (size_t i = 0; i < E; ++i)
            for (uint32_t j = enStart[i]; j < enStart[i + 1]; ++j) {
                bool dup = false;
                for (uint32_t j2 = enStart[i]; j2 < j; ++j2) dup = dup || en[j2] == en[j];
                if (!dup) ++incStart[en[j] + 1];
            }
        for (int v = 0; v < N; ++v) incStart[v + 1] += incStart[v];
        inc.assign(incStart[N], 0);
        {
            std::vector<uint32_t> fill(incStart.begin(), incStart.end() - 1);
            for (size_t i = 0; i < E; ++i)
                for (uint32_t j = enStart[i]; j < enStart[i + 1]; ++j) {
                    bool dup = false;
                    for (uint32_t j2 = enStart[i]; j2 < j; ++j2) dup = dup || en[j2] == en[j];
                    if (!dup) inc[fill[en[j]]++] = (uint32_t)i;
                }
        }
        std::vector<uint32_t> outCount((size_t)N, 0);
        std::vector<double> nodeBytes((size_t)N, 0.0);
        // Constants of an instance with only nodal blocks (M (x) I_dofs) repeat in every dof
        // column block of a tile (same values, same order): a tile then stores one period.
        const double constPeriod = (I.cRow.empty() && D > 1) ? (double)D : 1.0;
        for (int b = 0; b < N; ++b) {
            double bytes = 8.0 * D;
            for (int d = 0; d < D; ++d) {
                const int c = b + d * N;
                outCount[b] += (uint32_t)(colptr[c + 1] - colptr[c]);
                for (int q = colptr[c]; q < colptr[c + 1]; ++q)
                    bytes += 2.0 * (P.count[q] + (P.count[q] & 1u)) + (P.hasConst[q] ? 8.0 / constPeriod : 0.0);
            }
            nodeBytes[b] = bytes + outCount[b] * (1.0 + 4.0 * FUSED_CHUNK_WORDS / 64.0);
        }
        // block sums: a node adds its runs, node record, blocks (record, chunk record share,
        // list entries) and output chunk records; block constants are read from global memory,
        // not staged.  The lists' padding to each
        // chunk's longest is not counted, so this is an estimate, not a bound; a tile whose
        // exact LDS passes the budget is grown again with one node less (a few in ten
        // thousand tiles on cloths).
        std::vector<uint32_t> nodeBlocks(blocks ? (size_t)N : 0, 0);
        if (blocks) {
            for (int b = 0; b < N; ++b) {
                const int c = b;
                double nblk = 0.0, entries = 0.0;
                for (int q = colptr[c]; q < colptr[c + 1] && P.pattern.inner()[q] < N; ++q) {
                    nblk += 1.0;
                    entries += (double)P.count[q];
                }
                nodeBytes[b] = 8.0 * Db + 8.0 + outCount[b] * (20.0 / 128.0) + nblk * (4.0 + 8.0 / 64.0) + 2.0 * entries;
                nodeBlocks[b] = (uint32_t)nblk;
            }
        }

        std::vector<char> assigned((size_t)N, 0);
        for (int b = 0; b < N; ++b) assigned[b] = outCount[b] == 0;
        std::vector<int64_t> stamp(E, -1);
        std::vector<int64_t> nodeStamp((size_t)N, -1), nodeStep((size_t)N, -1);
        std::vector<uint32_t> localNode((size_t)N, 0), tileNodeList, newNodes;
        int64_t step = 0;
        std::vector<uint32_t> local(E, 0);
        std::vector<std::vector<uint32_t>> buckets;
        std::vector<size_t> heads;
        InstanceTiles& out = tiles[k];
        int64_t tileId = 0;
        size_t cursor = 0;
        size_t nodeCap = std::numeric_limits<size_t>::max();  // owned nodes of the tile being grown
        std::vector<uint32_t> tNodes;
        std::vector<std::vector<uint32_t>> tElems(NT);

        auto cost = [&](uint32_t v, int64_t id) {
            uint32_t c = 0;
            for (uint32_t q = incStart[v]; q < incStart[v + 1]; ++q) c += stamp[inc[q]] != id;
            return c;
        };
        auto push = [&](uint32_t v, uint32_t c) {
            if (buckets.size() <= c) {
                buckets.resize(c + 1);
                heads.resize(c + 1, 0);
            }
            buckets[c].push_back(v);
        };

        for (;;) {
            while (cursor < (size_t)N && assigned[cursor]) ++cursor;
            if (cursor == (size_t)N) break;
            for (auto& b : buckets) b.clear();
            std::fill(heads.begin(), heads.end(), 0);
            tNodes.clear();
            for (auto& te : tElems) te.clear();
            size_t elemBytes = 0, outs = 0, valueBytes = 0, tileBlocks = 0;
            double recipeBytes = 0.0;
            tileNodeList.clear();
            push((uint32_t)cursor, cost((uint32_t)cursor, tileId));
            for (;;) {
                // pop the cheapest frontier node (FIFO within a cost)
                uint32_t v = 0;
                bool found = false;
                uint32_t popped = 0;
                for (size_t c = 0; c < buckets.size() && !found; ++c)
                    while (heads[c] < buckets[c].size()) {
                        v = buckets[c][heads[c]++];
                        if (assigned[v]) continue;
                        popped = (uint32_t)c;
                        found = true;
                        break;
                    }
                if (!found) {
                    // frontier exhausted (a pocket enclosed by earlier tiles): continue with the
                    // next unassigned node rather than closing a small tile
                    size_t next = cursor;
                    while (next < (size_t)N && assigned[next]) ++next;
                    if (next == (size_t)N) break;
                    push((uint32_t)next, cost((uint32_t)next, tileId));
                    continue;
                }
                const uint32_t cur = cost(v, tileId);
                if (cur != popped) {
                    push(v, cur);
                    continue;
                }
                // LDS the node adds: its new elements, the nodes they bring in, its outputs
                size_t newBytes = 0, newValueBytes = 0;
                newNodes.clear();
                ++step;
                for (uint32_t q = incStart[v]; q < incStart[v + 1]; ++q) {
                    const uint32_t i = inc[q];
                    if (stamp[i] == tileId) continue;
                    newBytes += L[elemT[e0 + i]].elementBytes();
                    newValueBytes += 8 * (size_t)L[elemT[e0 + i]].valueStride();
                    for (uint32_t j = enStart[i]; j < enStart[i + 1]; ++j) {
                        const uint32_t n = en[j];
                        if (nodeStamp[n] == tileId || nodeStep[n] == step) continue;
                        nodeStep[n] = step;
                        newNodes.push_back(n);
                    }
                }
                double need = (double)(elemBytes + newBytes + nodeLds * (tileNodeList.size() + newNodes.size()));
                // block sums: the tile's outputs overlay its values and their zeros when more
                // (plus a padding slot before each run: at most Db per owned node, one at the end)
                if (blocks)
                    need += std::max(0.0, 8.0 * (double)(outs + outCount[v] + (size_t)Db * (tNodes.size() + 1) + 1) -
                                                           (double)(valueBytes + newValueBytes + 8 * (size_t)Db * Db));
                // output form: element values and the zero after them within 64 KiB (slots are u16
                // byte offsets); block form: values within 15-bit double indices
                const bool fits = need + recipeBytes + nodeBytes[v] <= ldsBudget &&
                                  (blocks ? valueBytes + newValueBytes + 8 * (size_t)Db * Db <= 8 * 0x7fff
                                          : valueBytes + newValueBytes + 8 <= 65536);
                const bool blockCap = blocks && tileBlocks + nodeBlocks[v] > FUSED_MAX_BLOCKS;
                if (!tNodes.empty() && (!fits || outs + outCount[v] > fo.maxOutputs || blockCap || tNodes.size() >= nodeCap))
                    break;
                if (!fits) throw std::invalid_argument("fused form: one node and its elements need more LDS than a tile holds");
                assigned[v] = 1;
                tNodes.push_back(v);
                outs += outCount[v];
                if (blocks) tileBlocks += nodeBlocks[v];
                elemBytes += newBytes;
                valueBytes += newValueBytes;
                recipeBytes += nodeBytes[v];
                for (uint32_t n : newNodes) {
                    nodeStamp[n] = tileId;
                    tileNodeList.push_back(n);
                }
                for (uint32_t q = incStart[v]; q < incStart[v + 1]; ++q) {
                    const uint32_t i = inc[q];
                    if (stamp[i] == tileId) continue;
                    stamp[i] = tileId;
                    tElems[(size_t)elemT[e0 + i]].push_back(i);
                    for (uint32_t j = enStart[i]; j < enStart[i + 1]; ++j)
                        if (!assigned[en[j]]) push(en[j], cost(en[j], tileId));
                }
            }

            // ---- emit the tile
            std::sort(tNodes.begin(), tNodes.end());
            std::sort(tileNodeList.begin(), tileNodeList.end());
            if (tileNodeList.size() > 65536) throw std::logic_error("fused form: tile nodes exceed 16-bit slots");
            for (size_t q = 0; q < tileNodeList.size(); ++q) localNode[tileNodeList[q]] = (uint32_t)q;
            std::vector<size_t> ldsBase(NT + 1, 0);
            for (size_t t = 0; t < NT; ++t) {
                std::sort(tElems[t].begin(), tElems[t].end());
                for (size_t q = 0; q < tElems[t].size(); ++q) local[tElems[t][q]] = (uint32_t)q;
                ldsBase[t + 1] = ldsBase[t] + tElems[t].size() * (size_t)L[t].valueStride();
            }
            // slot ldsBase[NT] is the tile's zero (padding of odd contribution lists); slots are
            // stored as LDS byte offsets (8 x index), so the values and the zero fit in 64 KiB
            if (!blocks && 8 * ldsBase[NT] >= 65536) throw std::logic_error("fused form: tile values exceed 64 KiB of byte offsets");
            const size_t constsBefore = out.consts.size();
            std::vector<uint32_t>& pool = out.pool;
            const size_t H = pool.size();
            out.header.push_back(H);
            pool.resize(H + HEADER + 3 * NT, 0);
            auto appendWords = [&](const void* data, size_t bytes) {
                const size_t off = pool.size() - H;
                const size_t words = (bytes + 3) / 4;
                pool.resize(pool.size() + words, 0);
                if (bytes) std::memcpy(&pool[H + off], data, bytes);
                return u32(off, "tile size");
            };
            size_t m = 0, stored = 0, contrib = 0, valueRegion = ldsBase[NT] + 1;
            if (blocks) {
                // ---- block sums: outputs in run order (dof-major columns of the owned nodes)
                const size_t nvals = ldsBase[NT], Z = (size_t)Db * Db, nOwned = tNodes.size();
                if (nvals + Z > 0x7fff) throw NotBlockStructured("tile values exceed 15-bit value indices");
                if (nOwned > 0x3fff) throw NotBlockStructured("more than 16383 owned nodes in a tile");
                const uint16_t zeroEntry = (uint16_t)nvals;
                // output slots of the tile in run order; every run starts at an even slot (a
                // padding slot before it when needed) and the slot count is even, so the outputs
                // are stored in pairs of slots
                std::vector<uint32_t> runs;  // (position, first slot)
                std::vector<uint32_t> colStart((size_t)Db * nOwned, 0);
                std::vector<char> padSlot;
                size_t sl = 0;
                uint32_t prevEnd = NONE;
                for (int d = 0; d < D; ++d)
                    for (size_t ni = 0; ni < nOwned; ++ni) {
                        const int c = (int)tNodes[ni] + d * N;
                        if (d >= Db) {
                            if (colptr[c + 1] != colptr[c]) throw NotBlockStructured("output column beyond the block dofs");
                            continue;
                        }
                        for (int q = colptr[c]; q < colptr[c + 1]; ++q, ++m, ++sl) {
                            if ((uint32_t)q != prevEnd) {
                                if (sl % 2) {
                                    padSlot.resize(sl + 1, 0);
                                    padSlot[sl++] = 1;
                                }
                                runs.push_back((uint32_t)q);
                                runs.push_back((uint32_t)sl);
                            }
                            if (q == colptr[c]) colStart[(size_t)d * nOwned + ni] = (uint32_t)sl;
                            prevEnd = (uint32_t)q + 1;
                        }
                    }
                if (sl % 2) {
                    padSlot.resize(sl + 1, 0);
                    padSlot[sl++] = 1;
                }
                padSlot.resize(sl, 0);
                const size_t m2 = sl;
                runs.push_back(0);
                runs.push_back((uint32_t)m2);
                if (m2 > 65536) throw NotBlockStructured("more than 65536 output slots in a tile");

                // blocks of every owned node b: list entries from its elements, E(b) (nodes
                // sharing an element), K(b) (the other dof-0 rows of column (b, 0)); the pattern's
                // columns, counts and constants are checked against the block structure
                struct Blk {
                    uint32_t ni, kind, rE, rEK, first, len;
                    bool hasConst;
                    double cval;
                };
                std::vector<Blk> blks;
                std::vector<uint16_t> ents;
                std::vector<uint32_t> nodeRec(2 * nOwned, 0);
                std::vector<std::pair<uint32_t, uint16_t>> pe;  // (node a, entry) of node b's elements
                std::vector<uint32_t> Eb, EK;
                const std::vector<int>& inner = P.pattern.inner();
                for (size_t ni = 0; ni < nOwned; ++ni) {
                    const uint32_t b = tNodes[ni];
                    pe.clear();
                    for (uint32_t qi = incStart[b]; qi < incStart[b + 1]; ++qi) {
                        const uint32_t i = inc[qi];
                        if (stamp[i] != tileId) throw std::logic_error("fused form: element of an owned node outside the tile");
                        const size_t t = (size_t)elemT[e0 + i];
                        const TemplateLayout& tl = L[t];
                        const size_t n = tl.diffNode.size();
                        size_t qb = n;
                        for (size_t p = 0; p < n; ++p)
                            if (en[enStart[i] + (uint32_t)tl.diffNode[p]] == b) qb = p;
                        if (qb == n) continue;  // b only read, not differentiated
                        const size_t vbase = ldsBase[t] + (size_t)local[i] * (size_t)tl.valueStride();
                        for (size_t p = 0; p < n; ++p) {
                            const size_t lo = std::min(p, qb), hi = std::max(p, qb);
                            const size_t v = vbase + (size_t)tl.blockBase[lo * n + hi];
                            pe.emplace_back(en[enStart[i] + (uint32_t)tl.diffNode[p]], (uint16_t)(v | (p > qb ? 0x8000u : 0u)));
                        }
                    }
                    std::stable_sort(pe.begin(), pe.end(), [](const std::pair<uint32_t, uint16_t>& x, const std::pair<uint32_t, uint16_t>& y) {
                        return x.first < y.first;
                    });
                    Eb.clear();
                    for (size_t j = 0; j < pe.size(); ++j)
                        if (j == 0 || pe[j].first != pe[j - 1].first) Eb.push_back(pe[j].first);
                    EK.clear();
                    for (int q = colptr[b]; q < colptr[b + 1] && inner[q] < N; ++q) EK.push_back((uint32_t)inner[q]);
                    const size_t nE = Eb.size();
                    if (EK.size() < nE || !std::includes(EK.begin(), EK.end(), Eb.begin(), Eb.end()))
                        throw NotBlockStructured("element pair without an output");
                    const size_t nK = EK.size() - nE;
                    if (EK.size() > 127) throw NotBlockStructured("node with more than 127 blocks");
                    // column (b, db): segment da holds E(b), and K(b) too when da == db
                    for (int db = 0; db < Db; ++db) {
                        const int c = (int)b + db * N;
                        int q = colptr[c];
                        for (int da = 0; da < Db; ++da)
                            for (uint32_t a : (da == db ? EK : Eb)) {
                                if (q >= colptr[c + 1] || inner[q] != (int)a + da * N) throw NotBlockStructured("output pattern");
                                ++q;
                            }
                        if (q != colptr[c + 1]) throw NotBlockStructured("output pattern");
                    }
                    nodeRec[2 * ni] = colStart[ni] | (Db > 1 ? colStart[nOwned + ni] << 16 : 0u);
                    nodeRec[2 * ni + 1] = (Db > 2 ? colStart[2 * nOwned + ni] : 0u) | (uint32_t)nE << 16 | (uint32_t)nK << 24;
                    size_t je = 0, jp = 0;
                    for (size_t r = 0; r < EK.size(); ++r) {
                        const uint32_t a = EK[r];
                        Blk B;
                        B.ni = (uint32_t)ni;
                        B.rEK = (uint32_t)r;
                        const bool inE = je < nE && Eb[je] == a;
                        B.rE = inE ? (uint32_t)je : 0u;
                        B.kind = inE ? (a == b ? 0u : 1u) : 2u;
                        B.first = (uint32_t)ents.size();
                        if (inE) {
                            ++je;
                            while (jp < pe.size() && pe[jp].first == a) ents.push_back(pe[jp++].second);
                        }
                        B.len = (uint32_t)(ents.size() - B.first);
                        // every output of the block: its contribution count and constant
                        const int q00 = colptr[b] + (int)r;  // entry (0, 0)
                        B.hasConst = P.hasConst[(size_t)q00] != 0;
                        B.cval = P.cval[(size_t)q00];
                        for (int db = 0; db < Db; ++db)
                            for (int da = 0; da < Db; ++da) {
                                if (!inE && da != db) continue;
                                const size_t q = (size_t)colptr[(int)b + db * N] + (size_t)da * nE + (db < da ? nK : 0) +
                                                 (da == db ? r : B.rE);
                                const bool hc = P.hasConst[q] != 0;
                                if (P.count[q] != B.len || hc != (da == db && B.hasConst) || (hc && P.cval[q] != B.cval))
                                    throw NotBlockStructured("output contributions");
                            }
                        if (B.len > 255) throw NotBlockStructured("block list longer than 255");
                        blks.push_back(B);
                    }
                }
                // blocks by kind, longest lists first: a chunk's lanes have near-equal lists
                const size_t nblk = blks.size(), noch = (m2 + 127) / 128;
                std::vector<uint32_t> order(nblk);
                std::iota(order.begin(), order.end(), 0u);
                std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
                    return blks[x].kind != blks[y].kind ? blks[x].kind < blks[y].kind : blks[x].len > blks[y].len;
                });
                // diagonal blocks whose lists are at least twice the longest full block's (a node's
                // every element against the elements of one edge) are summed by two lanes each,
                // l and l + 32 of a halved chunk, half of the list each; without this the chunk of
                // diagonal blocks takes its wave several times as long as any other
                size_t nd = 0;
                uint32_t Ld = 0, Lf = 0;
                for (const Blk& B : blks) {
                    if (B.kind == 0) {
                        ++nd;
                        Ld = std::max(Ld, B.len);
                    } else if (B.kind == 1) {
                        Lf = std::max(Lf, B.len);
                    }
                }
                size_t nhalf = Ld >= 4 && Ld >= 2 * Lf ? (nd + 31) / 32 : 0;
                if (64 * nhalf + nblk - nd > FUSED_MAX_BLOCKS) nhalf = 0;
                const size_t nsplit = nhalf ? nd : 0;  // blocks in halved chunks
                const size_t nb = 64 * nhalf + nblk - nsplit, nbch = (nb + 63) / 64;
                std::vector<uint32_t> blockRec(nb, 0xc000u), blockChunk(2 * nbch), lists;
                std::vector<double> cvals;
                for (size_t ch = 0; ch < nbch; ++ch) {
                    const bool halved = ch < nhalf;
                    // lane l: its block and the part [lo, hi) of the block's list it sums
                    struct LanePart {
                        const Blk* B;
                        uint32_t lo, hi;
                        bool partner;
                    };
                    LanePart lp[64];
                    uint32_t Lc = 0;
                    for (size_t l = 0; l < 64; ++l) {
                        lp[l] = {nullptr, 0, 0, false};
                        if (halved) {
                            const size_t k = ch * 32 + (l & 31);
                            if (k < nsplit) {
                                const Blk* B = &blks[order[k]];
                                const uint32_t h = (B->len + 1) / 2;
                                lp[l] = l < 32 ? LanePart{B, 0, h, false} : LanePart{B, h, B->len, true};
                            }
                        } else {
                            const size_t k = nsplit + (ch - nhalf) * 64 + l;
                            if (k < nblk) lp[l] = {&blks[order[k]], 0, blks[order[k]].len, false};
                        }
                        Lc = std::max(Lc, lp[l].hi - lp[l].lo);
                    }
                    const size_t pairs = (Lc + 1) / 2, lb = lists.size();
                    if (lb >= (1u << 24)) throw NotBlockStructured("block lists exceed 24-bit offsets");
                    blockChunk[2 * ch] = (uint32_t)lb | Lc << 24;
                    blockChunk[2 * ch + 1] = (uint32_t)cvals.size() | (halved ? 1u << 31 : 0u);
                    lists.resize(lb + 64 * pairs, 0);
                    for (size_t l = 0; l < 64; ++l) {
                        const LanePart& q = lp[l];
                        auto entry = [&](size_t j) -> uint32_t { return q.B && q.lo + j < q.hi ? ents[q.B->first + q.lo + j] : zeroEntry; };
                        for (size_t j = 0; j < pairs; ++j) lists[lb + 64 * j + l] = entry(2 * j) | entry(2 * j + 1) << 16;
                        if (!q.B) continue;
                        const Blk* B = q.B;
                        // a partner lane sums the second half of a diagonal block; its record
                        // has no constant, and it writes no outputs
                        const bool hc = B->hasConst && !q.partner;
                        blockRec[ch * 64 + l] = B->ni | B->kind << 14 | B->rE << 16 | B->rEK << 23 | (hc ? 1u << 30 : 0u);
                        if (hc) cvals.push_back(B->cval);
                        if (B->kind != 2 && !q.partner) contrib += (size_t)B->len * Z;
                    }
                }
                // output chunk records, per 128 slots (64 pairs): the run of the first pair, the
                // pairs after it where a run starts, the pairs whose second slot is padding
                std::vector<uint32_t> outChunk(5 * noch);
                for (size_t ch = 0, r = 0; ch < noch; ++ch) {
                    const size_t s0 = ch * 128;
                    while (runs[2 * (r + 1) + 1] <= s0) ++r;
                    uint64_t mask = 0, pad = 0;
                    for (size_t q = r + 1; runs[2 * q + 1] < std::min(m2, s0 + 128); ++q) mask |= 1ull << ((runs[2 * q + 1] - s0) / 2);
                    for (size_t i = 0; i < 64 && s0 + 2 * i < m2; ++i)
                        if (padSlot[s0 + 2 * i + 1]) pad |= 1ull << i;
                    outChunk[5 * ch] = (uint32_t)r;
                    outChunk[5 * ch + 1] = (uint32_t)mask;
                    outChunk[5 * ch + 2] = (uint32_t)(mask >> 32);
                    outChunk[5 * ch + 3] = (uint32_t)pad;
                    outChunk[5 * ch + 4] = (uint32_t)(pad >> 32);
                }
                const uint32_t recipeOff = appendWords(outChunk.data(), outChunk.size() * 4);
                appendWords(blockChunk.data(), blockChunk.size() * 4);
                appendWords(blockRec.data(), blockRec.size() * 4);
                const uint32_t nodeOff = appendWords(nodeRec.data(), nodeRec.size() * 4);
                const uint32_t runsOff = appendWords(runs.data(), runs.size() * 4);
                const uint32_t listOff = appendWords(lists.data(), lists.size() * 4);
                pool[H + 0] = u32(inOff[k], "input base");
                pool[H + 1] = (uint32_t)N;
                pool[H + 2] = (uint32_t)m2;
                pool[H + 3] = (uint32_t)nb;
                pool[H + 4] = recipeOff;
                pool[H + 5] = runsOff;
                pool[H + 6] = listOff;
                pool[H + 7] = nodeOff;
                if (!cvals.empty()) {
                    pool[H + 8] = u32(out.consts.size(), "constants");
                    out.consts.insert(out.consts.end(), cvals.begin(), cvals.end());
                } else {
                    pool[H + 8] = NONE;
                }
                pool[H + 9] = 0;  // result base, set at merge
                pool[H + 10] = listOff + (uint32_t)lists.size() - recipeOff;
                stored = cvals.size();
                if (stored >= (1u << 28)) throw std::length_error("fused form: too many output constants in a tile");
                pool[H + 11] = (uint32_t)stored;
                valueRegion = std::max(nvals + Z, m2);
                out.maxBlocks = std::max(out.maxBlocks, nb);
            } else {
                std::vector<uint32_t> runs;  // (position, first output)
                std::vector<uint8_t> counts;
                std::vector<uint16_t> slotList;
                std::vector<double> cvals;
                bool anyConst = false;
                m = 0;
                uint32_t prevEnd = NONE;
                for (int d = 0; d < D; ++d)
                    for (uint32_t b : tNodes) {
                        const int c = (int)b + d * N;
                        for (int q = colptr[c]; q < colptr[c + 1]; ++q, ++m) {
                            if ((uint32_t)q != prevEnd) {
                                runs.push_back((uint32_t)q);
                                runs.push_back((uint32_t)m);
                            }
                            prevEnd = (uint32_t)q + 1;
                            const uint32_t cnt = P.count[q];
                            if (cnt > 126) throw std::invalid_argument("fused form: an output has more than 126 element contributions");
                            // lists padded to even length with the zero slot (read in pairs)
                            counts.push_back((uint8_t)((cnt + (cnt & 1u)) | (P.hasConst[q] ? 0x80u : 0u)));
                            out.maxPairs = std::max(out.maxPairs, (int)((cnt + 1) / 2));
                            for (uint32_t r = 0; r < cnt; ++r) {
                                const uint64_t s = P.contrib[P.first[q] + r];
                                const size_t i = (size_t)(s >> 16) - e0;
                                const size_t slot = (size_t)(s & 0xffff);
                                if (stamp[i] != tileId) throw std::logic_error("fused form: contribution from an element outside the tile");
                                const size_t t = (size_t)elemT[e0 + i];
                                slotList.push_back((uint16_t)(8 * (ldsBase[t] + (size_t)local[i] * (size_t)L[t].valueStride() + slot)));
                            }
                            if (cnt & 1u) slotList.push_back((uint16_t)(8 * ldsBase[NT]));
                            if (P.hasConst[q]) {
                                cvals.push_back(P.cval[q]);
                                anyConst = true;
                            }
                        }
                    }
                runs.push_back(0);
                runs.push_back((uint32_t)m);
                const size_t nch = (m + 63) / 64;
                const size_t CW = FUSED_CHUNK_WORDS;
                std::vector<uint32_t> chunks(CW * nch);
                {
                    size_t sb = 0, cb = 0, r = 0;
                    for (size_t ch = 0; ch < nch; ++ch) {
                        const size_t o0 = ch * 64;
                        while (runs[2 * (r + 1) + 1] <= o0) ++r;
                        chunks[CW * ch] = (uint32_t)sb;
                        chunks[CW * ch + 1] = (uint32_t)r;
                        chunks[CW * ch + 2] = (uint32_t)cb;
                        uint64_t mask = 0;
                        for (size_t q = r + 1; runs[2 * q + 1] < std::min(m, o0 + 64); ++q) mask |= 1ull << (runs[2 * q + 1] - o0);
                        chunks[CW * ch + 3] = (uint32_t)mask;
                        chunks[CW * ch + 4] = (uint32_t)(mask >> 32);
                        for (size_t o = o0; o < std::min(m, o0 + 64); ++o) {
                            sb += counts[o] & 0x7fu;
                            cb += counts[o] >> 7;
                        }
                    }
                }

                const uint32_t runsOff = appendWords(runs.data(), runs.size() * 4);
                const uint32_t countsOff = appendWords(counts.data(), counts.size());
                const uint32_t slotsOff = appendWords(slotList.data(), slotList.size() * 2);
                const uint32_t chunksOff = appendWords(chunks.data(), chunks.size() * 4);
                pool[H + 0] = u32(inOff[k], "input base");
                pool[H + 1] = (uint32_t)N;
                pool[H + 2] = u32(m, "outputs per tile");
                pool[H + 3] = (uint32_t)nch;
                pool[H + 4] = runsOff;
                pool[H + 5] = countsOff;
                pool[H + 6] = slotsOff;
                pool[H + 7] = chunksOff;
                // periodic constants: the outputs split into D equal dof blocks (columns b + d N
                // for d = 0..D-1) and every block holds the same constant values in the same
                // order; then one block's constants are stored and H[11] carries the period
                uint32_t period = 0;
                stored = cvals.size();
                if (D > 1 && D < 16 && m % (size_t)D == 0 && cvals.size() % (size_t)D == 0 && !cvals.empty()) {
                    const size_t per = cvals.size() / (size_t)D, blk = m / (size_t)D;
                    bool ok = true;
                    size_t k = 0;
                    for (int d = 0; d < D && ok; ++d) {
                        size_t inBlock = 0;
                        for (size_t o = (size_t)d * blk; o < (size_t)(d + 1) * blk; ++o) inBlock += counts[o] >> 7;
                        ok = inBlock == per;
                        for (size_t j = 0; j < per && ok; ++j, ++k) ok = cvals[k] == cvals[j];
                    }
                    if (ok) {
                        period = (uint32_t)D;
                        stored = per;
                    }
                }
                if (anyConst) {
                    pool[H + 8] = u32(out.consts.size(), "constants");
                    out.consts.insert(out.consts.end(), cvals.begin(), cvals.begin() + (long)stored);
                } else {
                    pool[H + 8] = NONE;
                }
                pool[H + 9] = 0;  // result base, set at merge
                pool[H + 10] = chunksOff + (uint32_t)(CW * nch) - runsOff;
                if (stored >= (1u << 28)) throw std::length_error("fused form: too many output constants in a tile");
                pool[H + 11] = anyConst ? (uint32_t)stored | (period << 28) : 0u;
                contrib = slotList.size();
            }
            pool[H + 12] = 0;  // element image, set at merge
            pool[H + 13] = (uint32_t)tileNodeList.size();
            pool[H + 14] = appendWords(tileNodeList.data(), tileNodeList.size() * 4);
            size_t evals = 0;
            for (size_t t = 0; t < NT; ++t) {
                const auto& te = tElems[t];
                const TemplateLayout& tl = L[t];
                const size_t cnt = te.size();
                evals += cnt;
                std::vector<uint16_t> slotTab((size_t)tl.numNodes * cnt);
                for (size_t q = 0; q < cnt; ++q)
                    for (int j = 0; j < tl.numNodes; ++j)
                        slotTab[(size_t)j * cnt + q] = (uint16_t)localNode[en[enStart[te[q]] + (uint32_t)j]];
                const uint32_t slotOff = appendWords(slotTab.data(), slotTab.size() * 2);
                pool[H + HEADER + 3 * t] = (uint32_t)cnt;
                pool[H + HEADER + 3 * t + 1] = slotOff;
                pool[H + HEADER + 3 * t + 2] = u32(out.consts.size(), "constants");
                const size_t pb = out.consts.size();
                out.consts.resize(pb + (size_t)tl.numParams * cnt);
                for (size_t q = 0; q < cnt; ++q) {
                    const Arg* a = &args_[elemArg_[e0 + te[q]]];
                    for (size_t g = 0; g < tl.argParam.size(); ++g)
                        if (tl.argParam[g] >= 0) out.consts[pb + (size_t)tl.argParam[g] * cnt + q] = a[g].value;
                }
            }
            const size_t staged = fusedStagedDoubles(fp, &pool[H]);
            const size_t tileLds = 8 * (valueRegion + staged) + 8 * ((pool[H + 10] + 1) / 2) + (blocks ? 0 : 8 * stored);
            if (tileLds > fo.ldsBytes) {
                if (!blocks || tNodes.size() <= 1) throw std::logic_error("fused form: a tile passes the LDS budget");
                // the estimate let the tile grow past the budget: drop it, grow it again smaller
                pool.resize(H);
                out.header.pop_back();
                out.consts.resize(constsBefore);
                for (uint32_t v : tNodes) assigned[v] = 0;
                nodeCap = tNodes.size() - 1;
                ++tileId;
                continue;
            }
            nodeCap = std::numeric_limits<size_t>::max();
            out.cost.push_back((double)m + (double)ldsBase[NT]);
            out.values.push_back(ldsBase[NT]);
            out.stagedLds = std::max(out.stagedLds, 8 * staged);
            out.evaluations += evals;
            out.contributions += contrib;
            out.lds = std::max(out.lds, tileLds);
            out.elementLds = std::max(out.elementLds, 8 * (valueRegion + staged));
            out.outputs += m;
            ++tileId;
        }
        pats[k] = std::move(P.pattern);
    });
    timer("fused: patterns and tiles");

    // ---- merge instances (offsets first, then copies in parallel); order the tile list
    const size_t K = inst_.size();
    std::vector<size_t> poolBase(K), constBase(K), resBase(K), tileBase(K), imageBase(K);
    size_t nTiles = 0, pb = 0, cb = 0, rb = 0, ib = fp.numInputs;
    for (size_t k = 0; k < K; ++k) {
        tileBase[k] = nTiles;
        nTiles += tiles[k].header.size();
        imageBase[k] = ib;
        for (size_t v : tiles[k].values) ib += v;
        fp.stagedLdsBytes = std::max(fp.stagedLdsBytes, tiles[k].stagedLds);
    }
    u32(ib, "element images");
    fp.imageDoubles = ib - fp.numInputs;
    // pool: tile count, launch order, launch-order header table, then the tiles
    // (the node list stride rounded up to 64 words keeps the stride, a literal of the kernel
    // source, the same across most programs)
    size_t maxNodes = 0;
    for (size_t k = 0; k < K; ++k)
        for (size_t h : tiles[k].header) maxNodes = std::max<size_t>(maxNodes, tiles[k].pool[h + 13]);
    fp.headerNodes = (FUSED_HEADER + 3 * NT + 3) & ~size_t(3);
    fp.headerStride = fp.headerNodes + ((maxNodes + 63) & ~size_t(63));
    fp.headerTable = 1 + nTiles;
    pb = 1 + nTiles + nTiles * fp.headerStride;
    for (size_t k = 0; k < K; ++k) {
        if (tiles[k].outputs != pats[k].nonZeros())
            throw std::logic_error("fused form: tiles do not cover every output of an instance");
        poolBase[k] = pb;
        constBase[k] = cb;
        resBase[k] = rb;
        pb += tiles[k].pool.size();
        cb += tiles[k].consts.size();
        rb += pats[k].nonZeros();
        fp.elementEvaluations += tiles[k].evaluations;
        fp.numContributions += tiles[k].contributions;
        fp.ldsBytes = std::max(fp.ldsBytes, tiles[k].lds);
        fp.elementLdsBytes = std::max(fp.elementLdsBytes, tiles[k].elementLds);
        fp.maxPairs = std::max(fp.maxPairs, tiles[k].maxPairs);
        fp.maxBlocks = std::max(fp.maxBlocks, tiles[k].maxBlocks);
    }
    u32(pb, "tile pool");
    u32(cb, "constants");
    u32(rb, "results");
    fp.pool.assign(pb, 0);
    fp.consts.assign(cb, 0.0);
    fp.pool[0] = (uint32_t)nTiles;
    struct TileRef {
        uint32_t header;
        uint32_t inst, index;
        double cost;
    };
    std::vector<TileRef> refs(nTiles);
    parallelFor(K, [&](size_t k) {
        InstanceTiles& it = tiles[k];
        const uint32_t cbase = (uint32_t)constBase[k];
        size_t img = imageBase[k];
        for (size_t q = 0; q < it.header.size(); ++q) {
            uint32_t* H = &it.pool[it.header[q]];
            if (H[8] != NONE) H[8] += cbase;
            H[9] = (uint32_t)resBase[k];
            H[12] = (uint32_t)img;
            img += it.values[q];
            for (size_t t = 0; t < NT; ++t) H[HEADER + 3 * t + 2] += cbase;
            refs[tileBase[k] + q] = {(uint32_t)(poolBase[k] + it.header[q]), (uint32_t)k, (uint32_t)q, it.cost[q]};
        }
        std::copy(it.pool.begin(), it.pool.end(), fp.pool.begin() + (long)poolBase[k]);
        std::copy(it.consts.begin(), it.consts.end(), fp.consts.begin() + (long)constBase[k]);
        it = InstanceTiles();
    });
    patterns.clear();
    for (size_t k = 0; k < K; ++k) {
        const size_t nout = pats[k].nonZeros();
        fp.outStart.push_back(resBase[k]);
        fp.outLength.push_back(nout);
        io.outputSizes.push_back(nout);
        patterns.push_back(reportedPattern(k, std::move(pats[k])));
    }
    fp.numResults = rb;
    fp.tileHeader.resize(nTiles);
    fp.tileInstance.resize(nTiles);
    fp.tileIndex.resize(nTiles);
    fp.tileCost.resize(nTiles);
    for (size_t t = 0; t < nTiles; ++t) {
        fp.tileHeader[t] = refs[t].header;
        fp.tileInstance[t] = refs[t].inst;
        fp.tileIndex[t] = refs[t].index;
        fp.tileCost[t] = (float)refs[t].cost;
    }
    orderTiles(fp, fo.schedule);
    fp.numTiles = nTiles;
    if (fp.ldsBytes == 0) fp.ldsBytes = 8;
    timer("fused: merge");
    validateFused(fp);
    timer("fused: validate");
    return fp;
}

std::vector<float> tileStats(const FusedProgram& fp) {
    const size_t NT = fp.templates.size();
    const auto& P = fp.pool;
    std::vector<float> out(fp.numTiles * TILE_STATS);
    for (size_t t = 0; t < fp.numTiles; ++t) {
        const size_t H = P[1 + t];
        const uint32_t m = P[H + 2];
        const bool blockTile = fp.kind == 0 && fp.sums == FusedSums::Blocks;
        size_t elems = 0, staged = fp.kind == 1 ? P[H + 12] : 0, vals = 0, runs = 0, S = 0, pads = 0;
        for (size_t q = 0; q < NT; ++q) {
            const size_t cnt = P[H + HEADER + 3 * q];
            elems += cnt;
            vals += cnt * (size_t)fp.templates[q].valueStride;
        }
        if (blockTile) {
            // contributions: list entries of blocks with elements (not the zero entry), D^2 each
            const size_t nb = P[H + 3], nbch = (nb + 63) / 64, noch = ((size_t)m + 127) / 128;
            const size_t Z = (size_t)fp.blockDofs * fp.blockDofs;
            const size_t bch = H + P[H + 4] + 5 * noch, brec = bch + 2 * nbch, lst = H + P[H + 6];
            for (size_t ch = 0; ch < noch; ++ch)
                pads += (size_t)__builtin_popcount(P[H + P[H + 4] + 5 * ch + 3]) + (size_t)__builtin_popcount(P[H + P[H + 4] + 5 * ch + 4]);
            for (size_t ch = 0; ch < nbch; ++ch) {
                const size_t lb = P[bch + 2 * ch] & 0xffffffu, Lc = P[bch + 2 * ch] >> 24;
                for (size_t l = 0; l < 64 && ch * 64 + l < nb; ++l) {
                    if (((P[brec + ch * 64 + l] >> 14) & 3u) == 2) continue;
                    for (size_t j = 0; j < Lc; ++j) {
                        const uint32_t w = P[lst + lb + 64 * (j / 2) + l];
                        if ((j % 2 ? w >> 16 : w & 0xffffu) != vals) S += Z;
                    }
                }
            }
            while (P[H + P[H + 5] + 2 * runs + 1] != m) ++runs;
        } else {
            const uint8_t* cnts = reinterpret_cast<const uint8_t*>(&P[H + P[H + 5]]);
            if (fp.kind == 1)
                for (size_t o = 0; o < m; ++o) S += 2 * (cnts[o] & 0x7fu) + (cnts[o] >> 7);
            else
                for (size_t o = 0; o < m; ++o) S += cnts[o] & 0x7fu;
            while (P[H + P[H + 4] + 2 * runs + 1] != m) ++runs;
        }
        if (fp.kind == 0) staged = fusedStagedDoubles(fp, &P[H]);
        float* f = &out[t * TILE_STATS];
        f[0] = (float)(m - pads);
        f[1] = (float)S;
        f[2] = (float)elems;
        f[3] = (float)staged;
        f[4] = (float)vals;
        f[5] = (float)P[H + 10];
        f[6] = (float)(P[H + 11] & 0x0fffffffu);
        f[7] = (float)runs;
    }
    return out;
}

size_t fusedStagedDoubles(const FusedProgram& fp, const uint32_t* H) {
    size_t params = 0, words = 0;
    for (size_t t = 0; t < fp.templates.size(); ++t) {
        const size_t cnt = H[FUSED_HEADER + 3 * t];
        params += (size_t)fp.templates[t].numParams * cnt;
        words += ((size_t)fp.templates[t].numNodes * cnt + 1) / 2;
    }
    return params + (size_t)fp.stagedDofs * H[13] + (words + 1) / 2;
}

void orderTiles(FusedProgram& fp, FusedSchedule schedule) {
    const size_t n = fp.tileHeader.size();
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    switch (schedule) {
        case FusedSchedule::QueueLargest:
            std::stable_sort(order.begin(), order.end(),
                             [&](uint32_t a, uint32_t b) { return fp.tileCost[a] > fp.tileCost[b]; });
            break;
        case FusedSchedule::Interleaved:
        case FusedSchedule::Strided:
            std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
                return fp.tileIndex[a] != fp.tileIndex[b] ? fp.tileIndex[a] < fp.tileIndex[b]
                                                          : fp.tileInstance[a] < fp.tileInstance[b];
            });
            break;
        default:
            break;
    }
    for (size_t t = 0; t < n; ++t) fp.pool[1 + t] = fp.tileHeader[order[t]];
    if (fp.kind != 0) return;
    // element tiles: the headers again, in launch order at a fixed stride, with offsets made
    // absolute (pool indices), so a workgroup reads its header in one round of loads
    const size_t NT = fp.templates.size(), words = FUSED_HEADER + 3 * NT;
    if (fp.headerTable != 1 + n || fp.headerStride < words) throw std::logic_error("orderTiles: no header table in the pool");
    for (size_t t = 0; t < n; ++t) {
        const size_t H = fp.pool[1 + t], E = fp.headerTable + t * fp.headerStride;
        uint32_t* dst = &fp.pool[E];
        for (size_t w = 0; w < words; ++w) dst[w] = fusedHeaderOffset(w) ? (uint32_t)(H + fp.pool[H + w]) : fp.pool[H + w];
        const size_t nn = fp.pool[H + 13], tn = H + fp.pool[H + 14];
        if (fp.headerNodes < words || fp.headerNodes + nn > fp.headerStride) throw std::logic_error("orderTiles: node list exceeds the table stride");
        dst[14] = (uint32_t)(E + fp.headerNodes);
        for (size_t i = 0; i < fp.headerStride - fp.headerNodes; ++i) dst[fp.headerNodes + i] = i < nn ? fp.pool[tn + i] : 0u;
    }
}

namespace {

// Product-sum tiles (buildProductTiles): ranges, staged values, slots, runs, chunks.
void validateProducts(const FusedProgram& fp) {
    const auto& P = fp.pool;
    auto need = [](bool ok, const char* what) {
        if (!ok) throw std::logic_error(std::string("product tiles: invalid tile pool: ") + what);
    };
    need(!P.empty() && P[0] == fp.numTiles && P.size() >= 1 + fp.numTiles, "tile table");
    size_t totalContrib = 0;
    for (size_t t = 0; t < fp.numTiles; ++t) {
        const size_t H = P[1 + t];
        need(H + 14 <= P.size(), "header outside the pool");
        const uint32_t m = P[H + 2], nch = P[H + 3], nst = P[H + 12];
        need(nch == (m + 63) / 64 && P[H + 11] == 0 && P[H + 8] == 0xffffffffu, "header fields");
        const size_t runs = H + P[H + 4], counts = H + P[H + 5], slots = H + P[H + 6], chunks = H + P[H + 7],
                     staged = H + P[H + 13];
        need(P[H + 10] == P[H + 7] + 3 * nch - P[H + 4], "recipe tables not contiguous");
        need(staged + (size_t)nst <= P.size() && chunks + 3 * (size_t)nch <= P.size(), "tables outside the pool");
        need(nst <= 65536, "staged values exceed 16-bit slots");
        for (size_t s = 0; s < nst; ++s) need(P[staged + s] < fp.numInputs, "staged position outside the inputs");
        const uint8_t* cnts = reinterpret_cast<const uint8_t*>(&P[counts]);
        size_t S = 0;
        for (size_t o = 0; o < m; ++o) S += 2 * (cnts[o] & 0x7fu) + (cnts[o] >> 7);
        need(slots + (S + 1) / 2 <= P.size(), "slot list outside the pool");
        const uint16_t* sl = reinterpret_cast<const uint16_t*>(&P[slots]);
        for (size_t s = 0; s < S; ++s) need(sl[s] < nst, "slot outside the staged values");
        totalContrib += S;
        need(P[runs + 1] == 0 && P[runs + 3] == m && (size_t)P[H + 9] + P[runs] + m <= fp.numResults, "run");
        need(8 * (size_t)nst + 8 * (((size_t)P[H + 10] + 1) / 2) <= fp.ldsBytes, "tile LDS exceeds the launch's LDS");
        size_t sb = 0;
        for (size_t ch = 0; ch < nch; ++ch) {
            need(P[chunks + 3 * ch] == sb && P[chunks + 3 * ch + 1] == 0, "chunk");
            for (size_t o = ch * 64; o < std::min<size_t>(m, ch * 64 + 64); ++o) sb += 2 * (cnts[o] & 0x7fu) + (cnts[o] >> 7);
        }
    }
    need(totalContrib == fp.numContributions, "contribution count");
}

}  // namespace

namespace {

// Block-sum recipe of the element tile at H (layout in Fused.hpp); nvals: the tile's element
// values, staged: its staged doubles.  Walks every table as the kernel reads it and checks
// that the blocks' entries fill exactly the tile's output slots that are not padding, each
// once, and that the slot pairs store inside their runs.  Returns the tile's output
// contributions (list entries times D^2).
size_t validateBlockTile(const FusedProgram& fp, size_t H, size_t nvals, size_t staged) {
    const auto& P = fp.pool;
    auto need = [](bool ok, const char* what) {
        if (!ok) throw std::logic_error(std::string("fused form: invalid block-sum tile: ") + what);
    };
    const size_t D = (size_t)fp.blockDofs, Z = D * D, ND = D * (D + 1) / 2;
    need(D >= 1 && D <= 3, "block dofs");
    const size_t m2 = P[H + 2], nb = P[H + 3], noch = (m2 + 127) / 128, nbch = (nb + 63) / 64;
    need(m2 <= 65536 && m2 % 2 == 0 && nb <= fp.maxBlocks, "output slots or blocks");
    const size_t rec = H + P[H + 4], runs = H + P[H + 5], lst = H + P[H + 6], nod = H + P[H + 7];
    const size_t bch = rec + 5 * noch, brec = bch + 2 * nbch;
    need(brec + nb == nod && nod <= runs && runs < lst && (runs - nod) % 2 == 0 && (lst - runs) % 2 == 0,
         "recipe tables not contiguous");
    const size_t listWords = (size_t)P[H + 4] + P[H + 10] - P[H + 6];
    need(P[H + 4] + (size_t)P[H + 10] >= P[H + 6] && lst + listWords <= P.size(), "lists outside the pool");
    const size_t nOwned = (runs - nod) / 2, numRuns = (lst - runs) / 2 - 1;
    need(nOwned <= 0x3fff, "owned nodes");
    // runs: first slots even and increasing from 0 to the sentinel m2
    need(numRuns >= 1 || m2 == 0, "runs");
    need(P[runs + 2 * numRuns + 1] == m2 && (m2 == 0 || P[runs + 1] == 0), "run sentinel");
    for (size_t r = 0; r < numRuns; ++r) {
        const size_t first = P[runs + 2 * r + 1], next = P[runs + 2 * r + 3];
        need(first % 2 == 0 && next > first && next <= m2, "runs not increasing from even slots");
    }
    // node records: columns inside the slots
    for (size_t ni = 0; ni < nOwned; ++ni) {
        const uint32_t w0 = P[nod + 2 * ni], w1 = P[nod + 2 * ni + 1];
        const size_t cs[3] = {w0 & 0xffffu, w0 >> 16, w1 & 0xffffu}, nE = (w1 >> 16) & 0xffu, nK = w1 >> 24;
        need(nE + nK <= 127, "node blocks");
        for (size_t d = 0; d < 3; ++d) need(d < D ? cs[d] + D * nE + nK <= m2 : cs[d] == 0, "node columns");
    }
    // blocks: ranks inside their node's columns; their entries' slots, each once
    std::vector<char> seen(m2, 0);
    size_t numConst = 0, contrib = 0, listBase = 0;
    for (size_t ch = 0; ch < nbch; ++ch) {
        const uint32_t c0 = P[bch + 2 * ch], kw = P[bch + 2 * ch + 1];
        const size_t lb = c0 & 0xffffffu, Lc = c0 >> 24, pairs = (Lc + 1) / 2;
        const bool halved = (kw >> 31) != 0;
        need(lb == listBase && (kw & 0x7fffffffu) == numConst, "block chunk record");
        listBase += 64 * pairs;
        need(listBase <= listWords, "block lists");
        for (size_t l = 0; l < 64; ++l) {
            const size_t bi = ch * 64 + l;
            size_t kind = 3;
            // a halved chunk: lane l + 32 sums the rest of lane l's diagonal block (the same
            // record without a constant) and writes nothing; lanes without a block are 0xc000
            if (halved && l >= 32) need(bi < nb && P[brec + bi] == (P[brec + bi - 32] & ~(1u << 30)), "partner record");
            if (bi < nb && P[brec + bi] == 0xc000u) need(halved, "empty block record outside a halved chunk");
            if (bi < nb && P[brec + bi] != 0xc000u) {
                const uint32_t r = P[brec + bi];
                const size_t ni = r & 0x3fffu, rE = (r >> 16) & 0x7fu, rEK = (r >> 23) & 0x7fu;
                kind = (r >> 14) & 3u;
                need(ni < nOwned && kind <= 2 && (r >> 31) == 0, "block record");
                need(!halved || kind == 0, "halved chunk of other than diagonal blocks");
                numConst += (r >> 30) & 1u;
                const uint32_t w0 = P[nod + 2 * ni], w1 = P[nod + 2 * ni + 1];
                const size_t cs[3] = {w0 & 0xffffu, w0 >> 16, w1 & 0xffffu}, nE = (w1 >> 16) & 0xffu, nK = w1 >> 24;
                need(rEK < nE + nK && (kind == 2 || rE < nE), "block ranks");
                for (size_t db = 0; db < D && !(halved && l >= 32); ++db)
                    for (size_t da = 0; da < D; ++da) {
                        if (kind == 2 && da != db) continue;
                        const size_t pos = cs[db] + da * nE + (db < da ? nK : 0) + (da == db ? rEK : rE);
                        need(pos < m2 && !seen[pos], "block output slot");
                        seen[pos] = 1;
                    }
            }
            // entries a lane reads: every one of its chunk's list (kind 0: D(D+1)/2 values
            // each, kind 1: D^2); wings and lanes past the last block read none
            for (size_t j = 0; j < Lc; ++j) {
                const uint32_t w = P[lst + lb + 64 * (j / 2) + l];
                const uint32_t e = j % 2 ? w >> 16 : w & 0xffffu;
                const size_t v = e & 0x7fffu;
                if (kind == 0) need((e >> 15) == 0 && v + ND <= nvals + Z, "diagonal block entry");
                if (kind == 1) need(v + Z <= nvals + Z, "block entry");
                if (kind >= 2) need(e == nvals, "entry of a block without elements");
                if (kind <= 1 && e != nvals) contrib += Z;
            }
        }
    }
    need(listBase == listWords, "list words");
    // slot pairs: the run of each from the chunk records; a pair's second slot is padding
    // exactly when no block fills it, and a stored pair lies inside its run
    for (size_t ch = 0; ch < noch; ++ch) {
        const size_t r0 = P[rec + 5 * ch];
        const uint64_t mask = (uint64_t)P[rec + 5 * ch + 2] << 32 | P[rec + 5 * ch + 1];
        const uint64_t pad = (uint64_t)P[rec + 5 * ch + 4] << 32 | P[rec + 5 * ch + 3];
        need(r0 < numRuns && P[runs + 2 * r0 + 1] <= ch * 128 && P[runs + 2 * r0 + 3] > ch * 128, "output chunk run");
        for (size_t i = 0; i < 64; ++i) {
            const size_t s = ch * 128 + 2 * i;
            if (s >= m2) {
                need(!((mask >> i) & 1) && !((pad >> i) & 1), "chunk masks past the slots");
                continue;
            }
            const size_t r = r0 + (size_t)__builtin_popcountll(mask & (i == 63 ? ~0ull : ((2ull << i) - 1)));
            need(r < numRuns && P[runs + 2 * r + 1] <= s && s < P[runs + 2 * r + 3], "slot pair run");
            need(((mask >> i) & 1) == (i > 0 && P[runs + 2 * r + 1] == s), "run-start mask");
            const bool hiPad = (pad >> i) & 1;
            need(seen[s] && seen[s + 1] == !hiPad, "padding mask");
            need((size_t)P[H + 9] + P[runs + 2 * r] + (s - P[runs + 2 * r + 1]) + (hiPad ? 0 : 1) < fp.numResults,
                 "slot pair outside the results");
            need(hiPad || s + 1 < P[runs + 2 * r + 3], "slot pair across runs");
        }
    }
    const size_t stored = P[H + 11];
    need(stored == numConst, "block constant count");
    if (stored) need(P[H + 8] != 0xffffffffu && (size_t)P[H + 8] + stored <= fp.consts.size(), "block constants outside the constants");
    need(8 * (std::max(nvals + Z, m2) + staged) + 8 * (((size_t)P[H + 10] + 1) / 2) <= fp.ldsBytes,
         "tile LDS exceeds the launch's LDS");
    return contrib;
}

}  // namespace

void validateFused(const FusedProgram& fp) {
    if (fp.kind == 1) {
        validateProducts(fp);
        return;
    }
    const size_t NT = fp.templates.size();
    const auto& P = fp.pool;
    auto need = [](bool ok, const char* what) {
        if (!ok) throw std::logic_error(std::string("fused form: invalid tile pool: ") + what);
    };
    need(!P.empty() && P[0] == fp.numTiles && P.size() >= 1 + fp.numTiles, "tile table");
    size_t totalContrib = 0;
    for (size_t t = 0; t < fp.numTiles; ++t) {
        const size_t H = P[1 + t];
        need(H + HEADER + 3 * NT <= P.size(), "header outside the pool");
        const bool blockTile = fp.sums == FusedSums::Blocks;
        const uint32_t B = P[H], N = P[H + 1], m = P[H + 2], nch = blockTile ? (P[H + 2] + 63) / 64 : P[H + 3];
        need(nch == (m + 63) / 64, "chunk count");
        const size_t runs = H + P[H + 4], counts = H + P[H + 5], slots = H + P[H + 6], chunks = H + P[H + 7];
        const uint32_t cOff = P[H + 8], rbase = P[H + 9];
        const size_t CW = FUSED_CHUNK_WORDS;
        if (!blockTile) {
            need(chunks + CW * (size_t)nch <= P.size() && counts + ((size_t)m + 3) / 4 <= P.size(), "tables outside the pool");
            need(P[H + 10] == P[H + 7] + CW * nch - P[H + 4] && P[H + 4] <= P[H + 5] && P[H + 5] <= P[H + 6] &&
                     P[H + 6] <= P[H + 7],
                 "recipe tables not contiguous");
        }
        // LDS of the tile and element data
        size_t lds = 0, staged = 0;
        // tile nodes: sorted, inside the instance, every staged dof inside the inputs
        const size_t nn = P[H + 13], tn = H + P[H + 14];
        need(nn <= 65536 && tn + nn <= P.size(), "tile node list outside the pool");
        for (size_t i = 0; i < nn; ++i) need(P[tn + i] < N && (i == 0 || P[tn + i - 1] < P[tn + i]), "tile node list");
        need((size_t)B + (size_t)fp.stagedDofs * N <= fp.numInputs, "staged dofs outside the inputs");
        for (size_t q = 0; q < NT; ++q) {
            const FusedTemplate& ft = fp.templates[q];
            const size_t cnt = P[H + HEADER + 3 * q], tab = H + P[H + HEADER + 3 * q + 1], prm = P[H + HEADER + 3 * q + 2];
            need(tab + ((size_t)ft.numNodes * cnt + 1) / 2 <= P.size(), "slot table outside the pool");
            need(prm + (size_t)ft.numParams * cnt <= fp.consts.size(), "parameters outside the constants");
            const uint16_t* st = reinterpret_cast<const uint16_t*>(&P[tab]);
            for (size_t i = 0; i < (size_t)ft.numNodes * cnt; ++i) need(st[i] < nn, "slot outside the tile's nodes");
            for (size_t g = 0; g < ft.argDof.size(); ++g) need(ft.argDof[g] < fp.stagedDofs, "argument dof not staged");
            lds += cnt * (size_t)ft.valueStride;
        }
        staged = fusedStagedDoubles(fp, &P[H]);
        need(blockTile || 8 * (lds + 1) <= 65536, "tile values exceed 16-bit byte offsets");
        need((size_t)P[H + 12] >= fp.numInputs && (size_t)P[H + 12] + lds <= fp.numInputs + fp.imageDoubles,
             "element image outside the value buffer");
        need(8 * staged <= fp.stagedLdsBytes, "staged values exceed the element launch's LDS");
        if (blockTile) {
            totalContrib += validateBlockTile(fp, H, lds, staged);
            continue;
        }
        const uint8_t* cnts = reinterpret_cast<const uint8_t*>(&P[counts]);
        size_t S = 0, numConst = 0;
        for (size_t o = 0; o < m; ++o) {
            need((cnts[o] & 1u) == 0, "contribution list of odd length");
            need((int)(cnts[o] & 0x7fu) <= 2 * fp.maxPairs, "contribution list longer than the program's longest");
            S += cnts[o] & 0x7fu;
            numConst += cnts[o] >> 7;
        }
        // unrolled sums read maxPairs slot words from a lane's first one; the lane past the last
        // output starts at the end of the slot table, so the chunk records must follow it
        if (fp.maxPairs <= FUSED_UNROLLED_PAIRS && m)
            need(chunks == slots + S / 2 && CW * nch >= (size_t)fp.maxPairs, "slot words past the table's end");
        // stored constants: all of them, or one period of D equal dof blocks
        const size_t stored = P[H + 11] & 0x0fffffffu, period = P[H + 11] >> 28;
        if (period) {
            need(period > 1 && m % period == 0 && stored * period == numConst, "periodic output constants");
            const size_t blk = m / period;
            for (size_t d = 0; d < period; ++d) {
                size_t inBlock = 0;
                for (size_t o = d * blk; o < (d + 1) * blk; ++o) inBlock += cnts[o] >> 7;
                need(inBlock == stored, "periodic output constants per dof block");
            }
        } else {
            need(stored == numConst, "output constant count");
        }
        if (numConst) need(cOff != 0xffffffffu && (size_t)cOff + stored <= fp.consts.size(), "output constants outside the constants");
        need(8 * (lds + 1 + staged) + 8 * (((size_t)P[H + 10] + 1) / 2) + 8 * stored <= fp.ldsBytes,
             "tile LDS exceeds the launch's LDS");
        need(slots + (S + 1) / 2 <= P.size(), "slot list outside the pool");
        const uint16_t* sl = reinterpret_cast<const uint16_t*>(&P[slots]);
        // byte offsets of values; lds: the zero slot
        for (size_t s = 0; s < S; ++s) need(sl[s] % 8 == 0 && sl[s] / 8 <= lds, "contribution outside the tile's LDS");
        totalContrib += S;
        // runs: first outputs increasing from 0 to the sentinel m; positions inside the results
        size_t r = 0;
        need(m == 0 || P[runs + 1] == 0, "first run");
        while (P[runs + 2 * r + 1] != m) {
            need(runs + 2 * r + 3 < P.size(), "run list outside the pool");
            const size_t first = P[runs + 2 * r + 1], next = P[runs + 2 * r + 3];
            need(next > first && next <= m, "runs not increasing");
            need((size_t)rbase + P[runs + 2 * r] + (next - first) <= fp.numResults, "run outside the results");
            ++r;
        }
        const size_t numRuns = r;
        size_t sb = 0, cb = 0;
        for (size_t ch = 0; ch < nch; ++ch) {
            need(P[chunks + CW * ch] == sb && P[chunks + CW * ch + 2] == cb, "chunk slot or constant base");
            const size_t rr = P[chunks + CW * ch + 1];
            need(rr < numRuns && P[runs + 2 * rr + 1] <= ch * 64 && P[runs + 2 * rr + 3] > ch * 64, "chunk run");
            // run-start mask: exactly the runs that begin inside the chunk after its first output
            uint64_t mask = 0;
            for (size_t q = rr + 1; q < numRuns && P[runs + 2 * q + 1] < std::min<size_t>(m, ch * 64 + 64); ++q)
                mask |= 1ull << (P[runs + 2 * q + 1] - ch * 64);
            need(P[chunks + CW * ch + 3] == (uint32_t)mask && P[chunks + CW * ch + 4] == (uint32_t)(mask >> 32), "chunk run mask");
            for (size_t o = ch * 64; o < std::min<size_t>(m, ch * 64 + 64); ++o) {
                sb += cnts[o] & 0x7fu;
                cb += cnts[o] >> 7;
            }
        }
    }
    need(totalContrib == fp.numContributions, "contribution count");
    // the launch-order header table
    const size_t words = FUSED_HEADER + 3 * NT;
    need(fp.headerStride >= words && fp.headerTable + fp.numTiles * fp.headerStride <= P.size(), "header table outside the pool");
    need(fp.headerNodes >= words && fp.headerNodes <= fp.headerStride, "header table node lists");
    for (size_t t = 0; t < fp.numTiles; ++t) {
        const size_t H = P[1 + t], E = fp.headerTable + t * fp.headerStride;
        for (size_t w = 0; w < words; ++w)
            if (w != 14) need(P[E + w] == (fusedHeaderOffset(w) ? H + P[H + w] : P[H + w]), "header table entry");
        // the node list, copied after the entry's header words
        const size_t nn = P[H + 13], tn = H + P[H + 14];
        need(P[E + 14] == E + fp.headerNodes && fp.headerNodes + nn <= fp.headerStride, "header table node list");
        for (size_t i = 0; i < nn; ++i) need(P[E + fp.headerNodes + i] == P[tn + i], "header table node list copy");
    }
}

}  // namespace Sym
