#include "Program.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <set>
#include <stdexcept>

#include "Decomposition.hpp"
#include "Timer.hpp"

namespace Sym {

namespace {

// Blocks of equal shape at equal height form one group, the instances of one kernel (the
// role of src/ComputeUnit.cpp:33-56); blocks of equal height never read each other.
std::vector<std::vector<int>> group(const std::vector<DecomposedBlock>& blocks) {
    std::unordered_map<hash_t, int> groupMap;
    std::vector<std::vector<int>> ret;
    for (int i = 0; i < (int)blocks.size(); ++i) {
        const hash_t h = hash(blocks[i].shape, (long long)blocks[i].height);
        int& id = groupMap[h];
        if (id) ret[id - 1].push_back(i);
        else {
            ret.push_back({i});
            id = (int)ret.size();
        }
    }
    return ret;
}

inline long long varKey(const std::array<int, 2>& v) { return ((long long)v[1] << 32) ^ (unsigned)v[0]; }

}  // namespace

Program buildProgram(const std::vector<Symbolic>& outputs, const ProgramInputs& io, const BuildOptions& opt) {
    Program prog;
    prog.numInputs = io.inputDataSize;
    PhaseTimer timer;

    auto blocks = decompose(outputs, opt.decompositionThreshold, GLOBAL_INTERMEDIATE, true);
    timer("decompose");
    auto groupIds = group(blocks);
    std::vector<int> blockGroup(blocks.size());
    std::vector<KernelTemplate> kernels;
    kernels.reserve(groupIds.size());
    for (size_t g = 0; g < groupIds.size(); ++g) {
        std::vector<Symbolic> exprs(groupIds[g].size());
        for (size_t j = 0; j < groupIds[g].size(); ++j) {
            exprs[j] = blocks[groupIds[g][j]].code;
            blockGroup[groupIds[g][j]] = (int)g;
        }
        kernels.push_back(buildKernelTemplate(exprs, opt.simplifyTemplates));
    }
    timer("group + kernel templates");

    // dependencies: the group computing a child must run before its readers
    const size_t nk = kernels.size();
    std::vector<std::set<int>> readers(nk);
    for (size_t i = 0; i < blocks.size(); ++i)
        for (int j : blocks[i].reads)
            if (j != -1 && blockGroup[j] != blockGroup[i]) readers[blockGroup[j]].insert(blockGroup[i]);
    // kernels in dependency order (a kernel after every kernel it reads), each in the first
    // wave after all of those
    std::vector<int> pending(nk, 0), order;
    for (const auto& rs : readers)
        for (int r : rs) ++pending[(size_t)r];
    for (size_t k = 0; k < nk; ++k)
        if (!pending[k]) order.push_back((int)k);
    for (size_t q = 0; q < order.size(); ++q)
        for (int r : readers[(size_t)order[q]])
            if (--pending[(size_t)r] == 0) order.push_back(r);
    if (order.size() != nk) throw std::logic_error("buildProgram: cyclic kernel dependencies");

    std::vector<int> wave(nk, 0);
    for (int k : order)
        for (int r : readers[k]) wave[r] = std::max(wave[r], wave[k] + 1);

    // layout: intermediates after the inputs, outputs in the caller's order
    size_t acc = 0;
    for (size_t k = 0; k < io.outputSizes.size(); ++k) {
        prog.outStart.push_back(acc);
        prog.outLength.push_back(io.outputSizes[k]);
        acc += io.outputSizes[k];
    }

    std::unordered_map<long long, size_t> intermediatePos;
    std::vector<uint8_t> written(acc, 0);
    prog.kernels.resize(nk);
    for (size_t r = 0; r < nk; ++r) {
        const int k = order[r];
        const KernelTemplate& K = kernels[k];
        ProgramKernel& P = prog.kernels[r];
        P.n = K.instances;
        P.wave = wave[k];
        P.group = k;
        P.code = K.code;
        P.numOutSlots = K.outputColumns.size();

        const auto& otab = K.outputColumns;
        bool anyUser = false, anyInter = false;
        for (auto& col : otab)
            for (auto& v : col) {
                if (v[1] <= -10) anyUser = true;
                else anyInter = true;
            }
        if (anyUser && anyInter) throw std::logic_error("buildProgram: kernel mixes outputs and intermediates");
        P.userOutputs = anyUser;
        if (anyUser) {
            P.outposOffset = prog.outpos.size();
            for (auto& col : otab)
                for (auto& v : col) {
                    const int container = -(10 + v[1]);
                    if (container < 0 || (size_t)container >= io.outputSizes.size() || v[0] < 0 ||
                        (size_t)v[0] >= io.outputSizes[container])
                        throw std::logic_error("buildProgram: output variable out of range");
                    const size_t pos = prog.outStart[container] + (size_t)v[0];
                    if (written[pos]) throw std::logic_error("buildProgram: output written twice");
                    written[pos] = 1;
                    prog.outpos.push_back((uint32_t)pos);
                }
        } else {
            P.interBase = prog.numIntermediates;
            size_t slot = 0;
            for (auto& col : otab) {
                for (size_t i = 0; i < col.size(); ++i)
                    if (col[i][0] != -1) intermediatePos[varKey(col[i])] = P.interBase + slot * P.n + i;
                ++slot;
            }
            prog.numIntermediates += slot * P.n;
        }
    }
    for (size_t pos = 0; pos < acc; ++pos)
        if (!written[pos]) throw std::logic_error("buildProgram: an output value is never computed");
    timer("layout");

    for (size_t r = 0; r < nk; ++r) {
        const KernelTemplate& K = kernels[order[r]];
        ProgramKernel& P = prog.kernels[r];
        P.constOffset = prog.consts.size();
        for (const auto& col : K.constantColumns) prog.consts.insert(prog.consts.end(), col.begin(), col.end());
        P.numConstSlots = K.constantColumns.size();

        P.indexOffset = prog.index.size();
        P.access = K.inputLeaves;
        P.numVarSlots = K.inputColumns.size();
        std::vector<std::array<int, 2>> vd;
        for (const auto& col : K.inputColumns) vd.insert(vd.end(), col.begin(), col.end());
        for (auto& x : vd) {
            size_t pos;
            if (x[1] == GLOBAL_INTERMEDIATE || x[1] == EXPLICIT_INTERMEDIATE) {
                auto it = intermediatePos.find(varKey(x));
                if (it == intermediatePos.end()) throw std::logic_error("buildProgram: intermediate read before it is defined");
                pos = prog.numInputs + it->second;
            } else {
                auto it = io.inputOffsets.find(x[1]);
                if (it == io.inputOffsets.end())
                    throw std::invalid_argument("buildProgram: expression reads variable group " + std::to_string(x[1]) +
                                                " which is not an input of the unit");
                const size_t len = io.inputLengths.at(x[1]);
                if (x[0] < 0 || (size_t)x[0] >= len)
                    throw std::invalid_argument("buildProgram: variable {" + std::to_string(x[0]) + ", " +
                                                std::to_string(x[1]) + "} lies outside its input container");
                pos = it->second + (size_t)x[0];
            }
            if (pos > std::numeric_limits<uint32_t>::max()) throw std::length_error("buildProgram: index exceeds 32 bits");
            prog.index.push_back((uint32_t)pos);
        }
    }
    for (auto& P : prog.kernels) prog.numWaves = std::max(prog.numWaves, P.wave + 1);
    timer("tables");
    return prog;
}

}  // namespace Sym
