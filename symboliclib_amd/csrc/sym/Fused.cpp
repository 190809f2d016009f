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
};

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
            size_t elemBytes = 0, outs = 0, valueBytes = 0;
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
                const double need = (double)(elemBytes + newBytes + nodeLds * (tileNodeList.size() + newNodes.size()));
                // element values and the zero after them within 64 KiB: slots are u16 byte offsets
                const bool fits = need + recipeBytes + nodeBytes[v] <= ldsBudget && valueBytes + newValueBytes + 8 <= 65536;
                if (!tNodes.empty() && (!fits || outs + outCount[v] > fo.maxOutputs)) break;
                if (!fits) throw std::invalid_argument("fused form: one node and its elements need more LDS than a tile holds");
                assigned[v] = 1;
                tNodes.push_back(v);
                outs += outCount[v];
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
            if (8 * ldsBase[NT] >= 65536) throw std::logic_error("fused form: tile values exceed 64 KiB of byte offsets");
            std::vector<uint32_t> runs;  // (position, first output)
            std::vector<uint8_t> counts;
            std::vector<uint16_t> slotList;
            std::vector<double> cvals;
            bool anyConst = false;
            size_t m = 0;
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
            size_t stored = cvals.size();
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
            const size_t tileLds = 8 * (ldsBase[NT] + 1 + staged) + 8 * ((pool[H + 10] + 1) / 2) + 8 * stored;
            out.cost.push_back((double)m + (double)ldsBase[NT]);
            out.values.push_back(ldsBase[NT]);
            out.stagedLds = std::max(out.stagedLds, 8 * staged);
            out.evaluations += evals;
            out.contributions += slotList.size();
            out.lds = std::max(out.lds, tileLds);
            out.elementLds = std::max(out.elementLds, 8 * (ldsBase[NT] + 1 + staged));
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
        const uint8_t* cnts = reinterpret_cast<const uint8_t*>(&P[H + P[H + 5]]);
        size_t S = 0;
        for (size_t o = 0; o < m; ++o) S += cnts[o] & 0x7fu;
        size_t elems = 0, staged = fp.kind == 1 ? P[H + 12] : 0, vals = 0, runs = 0;
        if (fp.kind == 1) S = 0;
        if (fp.kind == 1)
            for (size_t o = 0; o < m; ++o) S += 2 * (cnts[o] & 0x7fu) + (cnts[o] >> 7);
        for (size_t q = 0; q < NT; ++q) {
            const size_t cnt = P[H + HEADER + 3 * q];
            elems += cnt;
            vals += cnt * (size_t)fp.templates[q].valueStride;
        }
        if (fp.kind == 0) staged = fusedStagedDoubles(fp, &P[H]);
        while (P[H + P[H + 4] + 2 * runs + 1] != m) ++runs;
        float* f = &out[t * TILE_STATS];
        f[0] = (float)m;
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
        const uint32_t B = P[H], N = P[H + 1], m = P[H + 2], nch = P[H + 3];
        need(nch == (m + 63) / 64, "chunk count");
        const size_t runs = H + P[H + 4], counts = H + P[H + 5], slots = H + P[H + 6], chunks = H + P[H + 7];
        const uint32_t cOff = P[H + 8], rbase = P[H + 9];
        const size_t CW = FUSED_CHUNK_WORDS;
        need(chunks + CW * (size_t)nch <= P.size() && counts + ((size_t)m + 3) / 4 <= P.size(), "tables outside the pool");
        // LDS of the tile and element data
        size_t lds = 0, staged = 0;
        need(P[H + 10] == P[H + 7] + CW * nch - P[H + 4] && P[H + 4] <= P[H + 5] && P[H + 5] <= P[H + 6] &&
                 P[H + 6] <= P[H + 7],
             "recipe tables not contiguous");
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
        need(8 * (lds + 1) <= 65536, "tile values exceed 16-bit byte offsets");
        need((size_t)P[H + 12] >= fp.numInputs && (size_t)P[H + 12] + lds <= fp.numInputs + fp.imageDoubles,
             "element image outside the value buffer");
        need(8 * staged <= fp.stagedLdsBytes, "staged values exceed the element launch's LDS");
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
