#include "Tiling.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <sstream>
#include <stdexcept>

#include "Timer.hpp"
#include "Traverse.hpp"

namespace Sym {

namespace {

struct Inst {
    uint32_t kernel;
    uint32_t inst;
};

}  // namespace

TiledProgram buildTiles(const Program& prog, const TileOptions& opt) {
    TiledProgram tp;
    PhaseTimer timer;
    const size_t K = prog.kernels.size();
    const size_t budget = opt.ldsBytes / sizeof(double);

    // ---- instance numbering
    std::vector<size_t> interBase(K, 0), outBase(K, 0);
    size_t numInter = 0, numOut = 0;
    std::vector<std::pair<size_t, uint32_t>> interRanges;  // (first intermediate offset, kernel)
    for (size_t r = 0; r < K; ++r) {
        const auto& P = prog.kernels[r];
        if (P.userOutputs) {
            outBase[r] = numOut;
            numOut += P.n;
        } else {
            interBase[r] = numInter;
            numInter += P.n;
            interRanges.emplace_back(P.interBase, (uint32_t)r);
        }
    }
    std::sort(interRanges.begin(), interRanges.end());
    std::vector<Inst> interInst(numInter), outInst(numOut);
    for (size_t r = 0; r < K; ++r) {
        const auto& P = prog.kernels[r];
        for (size_t i = 0; i < P.n; ++i) {
            if (P.userOutputs) outInst[outBase[r] + i] = {(uint32_t)r, (uint32_t)i};
            else interInst[interBase[r] + i] = {(uint32_t)r, (uint32_t)i};
        }
    }

    // intermediate position -> (kernel, slot, instance)
    auto decode = [&](uint32_t pos, uint32_t& kr, uint32_t& slot, uint32_t& inst) {
        const size_t off = pos - prog.numInputs;
        auto it = std::upper_bound(interRanges.begin(), interRanges.end(), std::make_pair(off, (uint32_t)K));
        if (it == interRanges.begin()) throw std::logic_error("buildTiles: bad intermediate position");
        --it;
        const auto& P = prog.kernels[it->second];
        const size_t rel = off - P.interBase;
        if (rel >= P.numOutSlots * P.n) throw std::logic_error("buildTiles: intermediate position outside its kernel");
        kr = it->second;
        slot = (uint32_t)(rel / P.n);
        inst = (uint32_t)(rel % P.n);
    };

    // ---- direct producers (intermediate instance ids) of every instance
    auto producersOf = [&](uint32_t r, uint32_t i, std::vector<uint32_t>& out) {
        out.clear();
        const auto& P = prog.kernels[r];
        for (size_t s = 0; s < P.numVarSlots; ++s) {
            const uint32_t pos = prog.index[P.indexOffset + s * P.n + i];
            if (pos < prog.numInputs) continue;
            uint32_t kr, slot, inst;
            decode(pos, kr, slot, inst);
            out.push_back((uint32_t)(interBase[kr] + inst));
        }
        std::sort(out.begin(), out.end());
        out.erase(std::unique(out.begin(), out.end()), out.end());
    };

    std::vector<uint64_t> outProdStart(numOut + 1, 0), interProdStart(numInter + 1, 0);
    std::vector<uint32_t> outProd, interProd, tmp;
    for (size_t o = 0; o < numOut; ++o) {
        producersOf(outInst[o].kernel, outInst[o].inst, tmp);
        outProd.insert(outProd.end(), tmp.begin(), tmp.end());
        outProdStart[o + 1] = outProd.size();
    }
    for (size_t g = 0; g < numInter; ++g) {
        producersOf(interInst[g].kernel, interInst[g].inst, tmp);
        interProd.insert(interProd.end(), tmp.begin(), tmp.end());
        interProdStart[g + 1] = interProd.size();
    }
    // consumers (outputs reading an intermediate directly)
    std::vector<uint64_t> consStart(numInter + 1, 0);
    for (uint32_t g : outProd) ++consStart[g + 1];
    for (size_t g = 0; g < numInter; ++g) consStart[g + 1] += consStart[g];
    std::vector<uint32_t> cons(outProd.size());
    {
        std::vector<uint64_t> fill(consStart.begin(), consStart.end() - 1);
        for (size_t o = 0; o < numOut; ++o)
            for (uint64_t q = outProdStart[o]; q < outProdStart[o + 1]; ++q) cons[fill[outProd[q]]++] = (uint32_t)o;
    }
    std::vector<uint32_t> slotsOf(numInter);
    for (size_t g = 0; g < numInter; ++g) slotsOf[g] = (uint32_t)prog.kernels[interInst[g].kernel].numOutSlots;

    timer("tiles: producers");
    // ---- seed order: outputs by their first output position (follows the caller's order)
    std::vector<uint32_t> seedOrder(numOut);
    {
        std::vector<uint32_t> firstPos(numOut);
        for (size_t o = 0; o < numOut; ++o) {
            const auto& P = prog.kernels[outInst[o].kernel];
            uint32_t m = std::numeric_limits<uint32_t>::max();
            for (size_t a = 0; a < P.numOutSlots; ++a) m = std::min(m, prog.outpos[P.outposOffset + a * P.n + outInst[o].inst]);
            firstPos[o] = m;
            seedOrder[o] = (uint32_t)o;
        }
        std::stable_sort(seedOrder.begin(), seedOrder.end(), [&](uint32_t a, uint32_t b) { return firstPos[a] < firstPos[b]; });
    }

    // ---- greedy growth
    std::vector<int32_t> tileOfOut(numOut, -1);
    std::vector<int32_t> inTileStamp(numInter, -1), enqStamp(numOut, -1);
    std::vector<std::vector<uint32_t>> tileOuts, tileInters;
    std::vector<uint32_t> dfs, newNodes;
    constexpr size_t NB = 256;
    // FIFO buckets by cost: breadth-first growth keeps tiles compact (LIFO grows snakes)
    std::vector<std::vector<uint32_t>> buckets(NB);
    std::vector<size_t> head(NB, 0);
    auto bucketEmpty = [&](size_t b) { return head[b] >= buckets[b].size(); };
    auto bucketPop = [&](size_t b) {
        const uint32_t v = buckets[b][head[b]++];
        if (head[b] == buckets[b].size()) {
            buckets[b].clear();
            head[b] = 0;
        }
        return v;
    };

    // cost in doubles of adding output o to tile t; fills newNodes
    auto costOf = [&](uint32_t o, int32_t t) -> size_t {
        newNodes.clear();
        dfs.clear();
        for (uint64_t q = outProdStart[o]; q < outProdStart[o + 1]; ++q) dfs.push_back(outProd[q]);
        size_t cost = 0;
        while (!dfs.empty()) {
            const uint32_t g = dfs.back();
            dfs.pop_back();
            if (inTileStamp[g] == t) continue;
            if (std::find(newNodes.begin(), newNodes.end(), g) != newNodes.end()) continue;
            newNodes.push_back(g);
            cost += slotsOf[g];
            for (uint64_t q = interProdStart[g]; q < interProdStart[g + 1]; ++q) dfs.push_back(interProd[q]);
        }
        return cost;
    };

    size_t seedIdx = 0;
    std::vector<uint32_t> frontier;  // candidates left over by the previous tile
    while (true) {
        while (seedIdx < numOut && tileOfOut[seedOrder[seedIdx]] >= 0) ++seedIdx;
        if (seedIdx >= numOut) break;
        // continue next to the previous tile when possible, so tiles grow side by side
        uint32_t seed = seedOrder[seedIdx];
        while (!frontier.empty()) {
            const uint32_t f = frontier.back();
            frontier.pop_back();
            if (tileOfOut[f] < 0) {
                seed = f;
                break;
            }
        }
        const int32_t t = (int32_t)tileOuts.size();
        tileOuts.emplace_back();
        tileInters.emplace_back();
        auto& outs = tileOuts.back();
        auto& inters = tileInters.back();
        size_t used = 0;

        const size_t c0 = costOf(seed, t);
        if (c0 > budget) {
            tp.reason = "an output needs more intermediate values than fit in one tile's LDS";
            return tp;
        }
        buckets[std::min(c0, NB - 1)].push_back(seed);
        enqStamp[seed] = t;
        size_t lowest = std::min(c0, NB - 1);

        size_t fillIdx = seedIdx;  // next output in caller order, used when no neighbour is queued
        while (true) {
            while (lowest < NB && bucketEmpty(lowest)) ++lowest;
            uint32_t o;
            size_t c;
            if (lowest < NB) {
                o = bucketPop(lowest);
                if (tileOfOut[o] >= 0) continue;
                c = costOf(o, t);
                if (std::min(c, NB - 1) < lowest) {  // cheaper than its bucket says: requeue
                    buckets[std::min(c, NB - 1)].push_back(o);
                    lowest = std::min(c, NB - 1);
                    continue;
                }
            } else {
                while (fillIdx < numOut && tileOfOut[seedOrder[fillIdx]] >= 0) ++fillIdx;
                if (fillIdx >= numOut) break;
                o = seedOrder[fillIdx];
                c = costOf(o, t);
                enqStamp[o] = t;
            }
            if (used + c > budget || outs.size() >= opt.maxOutputs) {
                // the cheapest candidate does not fit: the tile is full
                frontier.clear();
                frontier.push_back(o);
                for (size_t b = NB; b-- > 0;)
                    for (size_t q = head[b]; q < buckets[b].size(); ++q) frontier.push_back(buckets[b][q]);
                std::reverse(frontier.begin(), frontier.end());  // cheapest last = popped first
                for (auto& b : buckets) b.clear();
                std::fill(head.begin(), head.end(), 0);
                break;
            }
            tileOfOut[o] = t;
            outs.push_back(o);
            used += c;
            const std::vector<uint32_t> added(newNodes);  // costOf below reuses newNodes
            for (uint32_t g : added) {
                inTileStamp[g] = t;
                inters.push_back(g);
            }
            for (uint32_t g : added)
                for (uint64_t q = consStart[g]; q < consStart[g + 1]; ++q) {
                    const uint32_t o2 = cons[q];
                    if (tileOfOut[o2] >= 0 || enqStamp[o2] == t) continue;
                    enqStamp[o2] = t;
                    const size_t c2 = std::min(costOf(o2, t), NB - 1);
                    buckets[c2].push_back(o2);
                    lowest = std::min(lowest, c2);
                }
            if (c == 0) {
                // outputs whose closure is already complete: nothing new to enqueue
            }
        }
        tp.tiledIntermediates += inters.size();
        if (std::getenv("SYMLIB_TILE_DEBUG") && t < 8)
            std::fprintf(stderr, "[symlib] tile %d: %zu outputs, %zu intermediate instances, %zu doubles\n", t,
                         outs.size(), inters.size(), used);
    }
    tp.uniqueIntermediates = 0;
    {
        std::vector<char> needed(numInter, 0);
        for (auto& v : tileInters)
            for (uint32_t g : v) needed[g] = 1;
        for (char c : needed) tp.uniqueIntermediates += c;
        tp.kernelUsed.assign(K, 0);
        for (size_t r = 0; r < K; ++r)
            if (prog.kernels[r].userOutputs) tp.kernelUsed[r] = 1;
        for (size_t g = 0; g < numInter; ++g)
            if (needed[g]) tp.kernelUsed[interInst[g].kernel] = 1;
    }
    tp.numOutputInstances = numOut;

    timer("tiles: growth");
    // ---- tables
    const size_t T = tileOuts.size();
    tp.numTiles = T;
    tp.colOffset.assign(K + 1, 0);
    for (size_t r = 0; r < K; ++r) tp.colOffset[r + 1] = tp.colOffset[r] + prog.kernels[r].numVarSlots;
    std::vector<uint8_t> seenInput(tp.colOffset[K], 0), seenLds(tp.colOffset[K], 0);

    std::vector<uint32_t>& pool = tp.pool;
    pool.assign(T + 1, 0);
    std::vector<std::vector<uint32_t>> perKernel(K);
    std::vector<uint32_t> ldsOff(K), localIdx(numInter, 0);
    for (size_t t = 0; t < T; ++t) {
        for (auto& v : perKernel) v.clear();
        for (uint32_t o : tileOuts[t]) perKernel[outInst[o].kernel].push_back(outInst[o].inst);
        for (uint32_t g : tileInters[t]) perKernel[interInst[g].kernel].push_back(interInst[g].inst);
        size_t lds = 0;
        for (size_t r = 0; r < K; ++r) {
            auto& v = perKernel[r];
            std::sort(v.begin(), v.end());
            ldsOff[r] = (uint32_t)lds;
            if (!prog.kernels[r].userOutputs) {
                for (size_t j = 0; j < v.size(); ++j) localIdx[interBase[r] + v[j]] = (uint32_t)j;
                lds += prog.kernels[r].numOutSlots * v.size();
            }
        }
        if (lds > budget) throw std::logic_error("buildTiles: tile exceeds its LDS budget");
        tp.ldsBytes = std::max(tp.ldsBytes, lds * sizeof(double));

        const size_t hdr = pool.size();
        pool[t] = (uint32_t)hdr;
        pool.resize(hdr + 3 * K, 0);
        for (size_t r = 0; r < K; ++r) {
            const auto& P = prog.kernels[r];
            const auto& v = perKernel[r];
            const size_t cnt = v.size();
            pool[hdr + 3 * r] = (uint32_t)cnt;
            pool[hdr + 3 * r + 2] = ldsOff[r];
            if (!cnt) continue;
            const size_t data = pool.size();
            pool[hdr + 3 * r + 1] = (uint32_t)data;
            for (size_t s = 0; s < P.numVarSlots; ++s)
                for (size_t j = 0; j < cnt; ++j) {
                    const uint32_t pos = prog.index[P.indexOffset + s * P.n + v[j]];
                    if (pos < prog.numInputs) {
                        if (pos & LDS_FLAG) throw std::length_error("buildTiles: input position exceeds 31 bits");
                        pool.push_back(pos);
                        seenInput[tp.colOffset[r] + s] = 1;
                    } else {
                        uint32_t kr, slot, inst;
                        decode(pos, kr, slot, inst);
                        const size_t cntK = perKernel[kr].size();
                        const uint32_t li = ldsOff[kr] + slot * (uint32_t)cntK + localIdx[interBase[kr] + inst];
                        pool.push_back(LDS_FLAG | li);
                        seenLds[tp.colOffset[r] + s] = 1;
                    }
                }
            if (P.numConstSlots)
                for (size_t j = 0; j < cnt; ++j) pool.push_back(v[j]);
            if (P.userOutputs)
                for (size_t a = 0; a < P.numOutSlots; ++a)
                    for (size_t j = 0; j < cnt; ++j) pool.push_back(prog.outpos[P.outposOffset + a * P.n + v[j]]);
            if (pool.size() > std::numeric_limits<uint32_t>::max())
                throw std::length_error("buildTiles: tile tables exceed 32-bit offsets");
        }
    }
    pool[T] = (uint32_t)pool.size();
    tp.colKind.assign(tp.colOffset[K], 0);
    for (size_t k = 0; k < tp.colKind.size(); ++k) tp.colKind[k] = seenLds[k] ? (seenInput[k] ? 2 : 1) : 0;
    timer("tiles: tables");
    validateTiles(prog, tp);
    timer("tiles: validate");
    tp.valid = true;
    return tp;
}

// Every device access the tiled kernel derives from the tables stays in bounds:
// inputs (with the correlated-column offsets the code adds), LDS slots of the tile, constant
// rows, output positions.  Throws on the first violation.
void validateTiles(const Program& prog, const TiledProgram& tp) {
    const auto& pool = tp.pool;
    const size_t K = prog.kernels.size();
    const size_t numResults = prog.numResults();
    const size_t ldsDoubles = tp.ldsBytes / sizeof(double);
    auto fail = [](const std::string& m) { throw std::logic_error("tile tables invalid: " + m); };
    if (pool.size() < tp.numTiles + 1) fail("pool too small");
    // largest positive / negative access offset per base column
    std::vector<long> maxOff(tp.colOffset[K], 0), minOff(tp.colOffset[K], 0);
    for (size_t r = 0; r < K; ++r)
        for (const auto& va : prog.kernels[r].access) {
            const size_t col = tp.colOffset[r] + va.column;
            maxOff[col] = std::max<long>(maxOff[col], va.offset);
            minOff[col] = std::min<long>(minOff[col], va.offset);
        }
    for (size_t t = 0; t < tp.numTiles; ++t) {
        const size_t hdr = pool[t];
        if (hdr + 3 * K > pool.size()) fail("header out of range");
        size_t lds = 0;
        for (size_t r = 0; r < K; ++r) {
            const auto& P = prog.kernels[r];
            const size_t cnt = pool[hdr + 3 * r], data = pool[hdr + 3 * r + 1], ldsOff = pool[hdr + 3 * r + 2];
            if (!cnt) continue;
            if (!tp.kernelUsed[r]) fail("instances of an unused kernel");
            const size_t rows = P.numVarSlots + (P.numConstSlots ? 1 : 0) + (P.userOutputs ? P.numOutSlots : 0);
            if (data + rows * cnt > pool.size()) fail("kernel data out of range");
            if (!P.userOutputs) {
                if (ldsOff + P.numOutSlots * cnt > ldsDoubles) fail("LDS region beyond the launch size");
                lds = std::max(lds, ldsOff + P.numOutSlots * cnt);
            }
            for (size_t s = 0; s < P.numVarSlots; ++s)
                for (size_t j = 0; j < cnt; ++j) {
                    const uint32_t v = pool[data + s * cnt + j];
                    const size_t col = tp.colOffset[r] + s;
                    if (v & LDS_FLAG) {
                        if ((v & ~LDS_FLAG) >= ldsDoubles) fail("LDS operand out of range");
                        if (maxOff[col] || minOff[col]) fail("offset access into LDS");
                    } else if ((long)v + minOff[col] < 0 || (size_t)((long)v + maxOff[col]) >= prog.numInputs) {
                        fail("input operand out of range");
                    }
                }
            size_t row = P.numVarSlots;
            if (P.numConstSlots) {
                for (size_t j = 0; j < cnt; ++j)
                    if (pool[data + row * cnt + j] >= P.n) fail("instance id out of range");
                if (P.constOffset + P.numConstSlots * P.n > prog.consts.size()) fail("constant table out of range");
                ++row;
            }
            if (P.userOutputs)
                for (size_t a = 0; a < P.numOutSlots; ++a)
                    for (size_t j = 0; j < cnt; ++j)
                        if (pool[data + (row + a) * cnt + j] >= numResults) fail("output position out of range");
        }
        if (lds > ldsDoubles) fail("tile LDS beyond the launch size");
    }
}

}  // namespace Sym
