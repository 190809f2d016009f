// Product-sum form of B = A^T A + A: pattern by Gustavson's row merge, tiles of consecutive
// output columns, staged values of A, per-output product lists (tile layout in
// Fused.hpp).
#include <algorithm>
#include <atomic>
#include <cstring>
#include <exception>
#include <mutex>
#include <thread>
#include <limits>
#include <numeric>
#include <stdexcept>

#include "Fused.hpp"
#include "Parallel.hpp"
#include "Timer.hpp"

namespace Sym {

namespace {

constexpr uint32_t NONE = 0xffffffffu;
constexpr int HEADER = 14;

}  // namespace

FusedProgram buildProductTiles(int n, const int* colptr, const int* rowidx, const FusedOptions& fo,
                               std::vector<int>& outer, std::vector<int>& inner) {
    if (fo.ldsBytes > FUSED_MAX_LDS_BYTES) throw std::invalid_argument("product tiles: more LDS per workgroup than a CU has (160 KiB)");
    PhaseTimer timer;
    const int T = fo.threadsPerBlock;
    if (T <= 0 || T > 1024 || T % 64) throw std::invalid_argument("product tiles: threads per block must be a multiple of 64 up to 1024");
    if (n <= 0) throw std::invalid_argument("product tiles: empty matrix");
    const size_t nnzA = (size_t)colptr[n];
    if (nnzA > std::numeric_limits<uint32_t>::max() / 2) throw std::length_error("product tiles: too many entries in A");

    // CSR of A with CSC positions
    std::vector<uint32_t> rowptr((size_t)n + 1, 0), rcol(nnzA), rpos(nnzA);
    for (size_t p = 0; p < nnzA; ++p) ++rowptr[(size_t)rowidx[p] + 1];
    for (int i = 0; i < n; ++i) rowptr[i + 1] += rowptr[i];
    {
        std::vector<uint32_t> fill(rowptr.begin(), rowptr.end() - 1);
        for (int j = 0; j < n; ++j)
            for (int p = colptr[j]; p < colptr[j + 1]; ++p) {
                const uint32_t q = fill[rowidx[p]]++;
                rcol[q] = (uint32_t)j;
                rpos[q] = (uint32_t)p;
            }
    }

    // pattern of B, column by column (rows of A(:, j) and every column meeting them)
    std::vector<std::vector<int>> colRows((size_t)n);
    {
        const unsigned W = workerThreads();
        std::vector<std::vector<int>> mark(W, std::vector<int>((size_t)n, -1));
        std::atomic<int> next{0};
        std::vector<std::thread> pool;
        for (unsigned w = 0; w < W; ++w)
            pool.emplace_back([&, w] {
                std::vector<int>& mk = mark[w];
                for (int j; (j = next.fetch_add(64)) < n;)
                    for (int jj = j; jj < std::min(n, j + 64); ++jj) {
                        std::vector<int>& rows = colRows[jj];
                        for (int p = colptr[jj]; p < colptr[jj + 1]; ++p) {
                            const int k = rowidx[p];
                            if (mk[k] != jj) {
                                mk[k] = jj;
                                rows.push_back(k);
                            }
                            for (uint32_t q = rowptr[k]; q < rowptr[k + 1]; ++q) {
                                const int i = (int)rcol[q];
                                if (mk[i] != jj) {
                                    mk[i] = jj;
                                    rows.push_back(i);
                                }
                            }
                        }
                        std::sort(rows.begin(), rows.end());
                    }
            });
        for (auto& th : pool) th.join();
    }
    outer.assign((size_t)n + 1, 0);
    for (int j = 0; j < n; ++j) outer[j + 1] = outer[j] + (int)colRows[j].size();
    if ((size_t)outer[n] > std::numeric_limits<uint32_t>::max() / 2) throw std::length_error("product tiles: too many outputs");
    inner.resize((size_t)outer[n]);
    for (int j = 0; j < n; ++j) std::copy(colRows[j].begin(), colRows[j].end(), inner.begin() + outer[j]);
    colRows = std::vector<std::vector<int>>();
    timer("products: pattern");

    // tiles of consecutive columns: staged positions = all entries of the rows of A met by
    // the tile's columns
    // LDS of a tile: 8 bytes per staged value plus its recipe (1 byte per output, 2 bytes per
    // slot, 12 bytes per 64-output chunk, 16 bytes of runs), within fo.ldsBytes
    std::vector<int> tileStart{0};
    {
        std::vector<int> mark((size_t)n, -1);
        size_t bytes = 16;
        int t = 0;
        size_t outs = 0;
        for (int j = 0; j < n; ++j) {
            size_t add = 0, prod = 0;
            for (int p = colptr[j]; p < colptr[j + 1]; ++p) {
                const int k = rowidx[p];
                const size_t r = rowptr[k + 1] - rowptr[k];
                prod += r;
                if (mark[k] != t) add += r;
            }
            const size_t no = (size_t)(outer[j + 1] - outer[j]);
            const size_t recipe = no + 2 * (2 * prod + (size_t)(colptr[j + 1] - colptr[j])) + (12 * no + 63) / 64 + 8;
            size_t need = 8 * add + recipe;
            if (j > tileStart.back() && (bytes + need > fo.ldsBytes || outs + no > fo.maxOutputs)) {
                tileStart.push_back(j);
                ++t;
                bytes = 16;
                outs = 0;
                add = 0;
                for (int p = colptr[j]; p < colptr[j + 1]; ++p) add += rowptr[rowidx[p] + 1] - rowptr[rowidx[p]];
                need = 8 * add + recipe;
            }
            if (bytes + need > fo.ldsBytes) throw std::invalid_argument("product tiles: one column needs more LDS than a tile holds");
            for (int p = colptr[j]; p < colptr[j + 1]; ++p) mark[rowidx[p]] = t;
            bytes += need;
            outs += no;
        }
        tileStart.push_back(n);
    }
    const size_t nTiles = tileStart.size() - 1;
    timer("products: tile boundaries");

    struct TileOut {
        std::vector<uint32_t> pool;
        size_t lds = 0, contributions = 0, staged = 0;
    };
    std::vector<TileOut> tiles(nTiles);
    {
        const unsigned W = workerThreads();
        std::vector<std::vector<int>> rowMark(W, std::vector<int>((size_t)n, -1));
        std::vector<std::vector<int32_t>> ldsOf(W, std::vector<int32_t>(nnzA, -1));
        std::vector<std::vector<int32_t>> outOf(W, std::vector<int32_t>((size_t)n, -1));
        std::atomic<size_t> next{0};
        std::exception_ptr err;
        std::mutex errMutex;
        std::vector<std::thread> pool;
        for (unsigned w = 0; w < W; ++w)
            pool.emplace_back([&, w] {
                try {
                    std::vector<int>& rmk = rowMark[w];
                    std::vector<int32_t>& lds = ldsOf[w];
                    std::vector<int32_t>& oix = outOf[w];
                    for (size_t t; (t = next.fetch_add(1)) < nTiles;) {
                        const int j0 = tileStart[t], j1 = tileStart[t + 1];
                        // staged positions: rows met, in position order
                        std::vector<uint32_t> pos;
                        for (int j = j0; j < j1; ++j)
                            for (int p = colptr[j]; p < colptr[j + 1]; ++p) {
                                const int k = rowidx[p];
                                if (rmk[k] != (int)t) {
                                    rmk[k] = (int)t;
                                    for (uint32_t q = rowptr[k]; q < rowptr[k + 1]; ++q) pos.push_back(rpos[q]);
                                }
                            }
                        std::sort(pos.begin(), pos.end());
                        if (pos.size() > 65536) throw std::logic_error("product tiles: staged values exceed 16-bit slots");
                        for (size_t s = 0; s < pos.size(); ++s) lds[pos[s]] = (int32_t)s;
                        // recipe: outputs in CSC order; per output the A(i, j) term and its products
                        const size_t m = (size_t)(outer[j1] - outer[j0]);
                        std::vector<uint8_t> counts(m, 0);
                        std::vector<std::vector<uint16_t>> lists(m);
                        for (int j = j0; j < j1; ++j) {
                            const size_t ob = (size_t)(outer[j] - outer[j0]);
                            for (int q = outer[j]; q < outer[j + 1]; ++q) oix[inner[q]] = (int32_t)(ob + (size_t)(q - outer[j]));
                            for (int p = colptr[j]; p < colptr[j + 1]; ++p) {
                                const int k = rowidx[p];
                                const uint16_t lkj = (uint16_t)lds[(size_t)p];
                                for (uint32_t q = rowptr[k]; q < rowptr[k + 1]; ++q) {
                                    auto& L = lists[(size_t)oix[rcol[q]]];
                                    L.push_back((uint16_t)lds[rpos[q]]);
                                    L.push_back(lkj);
                                }
                            }
                            // the A(i, j) term goes first in its output's list
                            for (int p = colptr[j]; p < colptr[j + 1]; ++p) {
                                const size_t o = (size_t)oix[rowidx[p]];
                                auto& L = lists[o];
                                L.insert(L.begin(), (uint16_t)lds[(size_t)p]);
                                counts[o] = 0x80u;
                            }
                        }
                        std::vector<uint16_t> slots;
                        for (size_t o = 0; o < m; ++o) {
                            const size_t pairs = (lists[o].size() - (counts[o] >> 7)) / 2;
                            if (pairs > 127) throw std::invalid_argument("product tiles: an output has more than 127 products");
                            counts[o] = (uint8_t)(counts[o] | pairs);
                            slots.insert(slots.end(), lists[o].begin(), lists[o].end());
                        }
                        for (uint32_t p : pos) lds[p] = -1;
                        const size_t nch = (m + 63) / 64;
                        std::vector<uint32_t> chunks(3 * nch);
                        {
                            size_t sb = 0;
                            for (size_t ch = 0; ch < nch; ++ch) {
                                chunks[3 * ch] = (uint32_t)sb;
                                chunks[3 * ch + 1] = 0;
                                chunks[3 * ch + 2] = 0;
                                for (size_t o = ch * 64; o < std::min(m, ch * 64 + 64); ++o)
                                    sb += 2 * (counts[o] & 0x7fu) + (counts[o] >> 7);
                            }
                        }
                        const uint32_t runs[4] = {(uint32_t)outer[j0], 0, 0, (uint32_t)m};
                        TileOut& to = tiles[t];
                        std::vector<uint32_t>& P = to.pool;
                        P.assign(HEADER, 0);
                        auto append = [&](const void* data, size_t bytes) {
                            const size_t off = P.size();
                            P.resize(off + (bytes + 3) / 4, 0);
                            if (bytes) std::memcpy(&P[off], data, bytes);
                            return (uint32_t)off;
                        };
                        const uint32_t runsOff = append(runs, sizeof(runs));
                        const uint32_t countsOff = append(counts.data(), counts.size());
                        const uint32_t slotsOff = append(slots.data(), slots.size() * 2);
                        const uint32_t chunksOff = append(chunks.data(), chunks.size() * 4);
                        const uint32_t posOff = append(pos.data(), pos.size() * 4);
                        P[0] = 0;
                        P[1] = 0;
                        P[2] = (uint32_t)m;
                        P[3] = (uint32_t)nch;
                        P[4] = runsOff;
                        P[5] = countsOff;
                        P[6] = slotsOff;
                        P[7] = chunksOff;
                        P[8] = NONE;
                        P[9] = 0;
                        P[10] = chunksOff + 3 * (uint32_t)nch - runsOff;
                        P[11] = 0;
                        P[12] = (uint32_t)pos.size();
                        P[13] = posOff;
                        to.lds = 8 * pos.size() + 8 * ((P[10] + 1) / 2);
                        to.contributions = slots.size();
                        to.staged = pos.size();
                    }
                } catch (...) {
                    std::lock_guard<std::mutex> g(errMutex);
                    if (!err) err = std::current_exception();
                    next = nTiles;
                }
            });
        for (auto& th : pool) th.join();
        if (err) std::rethrow_exception(err);
    }
    timer("products: tiles");

    FusedProgram fp;
    fp.kind = 1;
    fp.numInputs = nnzA;
    fp.numResults = (size_t)outer[n];
    fp.outStart = {0};
    fp.outLength = {fp.numResults};
    fp.numTiles = nTiles;
    fp.numElements = 0;
    size_t total = 1 + nTiles;
    for (auto& t : tiles) total += t.pool.size();
    if (total > std::numeric_limits<uint32_t>::max()) throw std::length_error("product tiles: tile pool exceeds 32 bits");
    fp.pool.assign(total, 0);
    fp.pool[0] = (uint32_t)nTiles;
    size_t off = 1 + nTiles;
    for (size_t t = 0; t < nTiles; ++t) {
        fp.tileHeader.push_back((uint32_t)off);
        fp.tileInstance.push_back(0);
        fp.tileIndex.push_back((uint32_t)t);
        fp.tileCost.push_back((float)(tiles[t].contributions + tiles[t].staged));
        std::copy(tiles[t].pool.begin(), tiles[t].pool.end(), fp.pool.begin() + (long)off);
        off += tiles[t].pool.size();
        fp.ldsBytes = std::max(fp.ldsBytes, tiles[t].lds);
        fp.elementLdsBytes = std::max(fp.elementLdsBytes, 8 * tiles[t].staged);
        fp.numContributions += tiles[t].contributions;
        fp.elementEvaluations += tiles[t].staged;
        tiles[t] = TileOut();
    }
    orderTiles(fp, fo.schedule);
    timer("products: merge");
    validateFused(fp);
    timer("products: validate");
    return fp;
}

}  // namespace Sym
