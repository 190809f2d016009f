// Compiled expression program: the kernels a ComputeUnit evaluates, their tables, and the
// memory layout they address.  Built from the decomposed blocks the same way as
// src/ComputeUnit.cpp:230-366 of the reference (group -> Kernel -> dependency order ->
// layout -> constants -> gather indices), with two layout changes for the GPU:
//   * kernels that do not depend on each other are put in the same wave and launched
//     together (one launch per wave instead of one per group);
//   * outputs are written straight to their final position in the caller's output order
//     through a placement table, instead of to scratch slots that a second kernel gathers.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "ComputeKernel.hpp"
#include "Symbolic.hpp"

namespace Sym {

struct ProgramKernel {
    size_t n = 0;          // instances
    int wave = 0;          // launch wave (0 runs first)
    Symbolic code;         // BLOCK template, leaves VAR(k,-2/-3/-4) and LOCAL_VAR
    bool userOutputs = false;
    size_t indexOffset = 0;   // first entry in Program::index
    size_t constOffset = 0;   // first entry in Program::consts
    size_t outposOffset = 0;  // first entry in Program::outpos (user outputs only)
    size_t interBase = 0;     // first intermediate slot (offset after the inputs) of this kernel
    size_t numVarSlots = 0, numConstSlots = 0, numOutSlots = 0;
    std::vector<ColumnRef> access;  // per input leaf of the template: stored column and offset
    int group = -1;
};

struct Program {
    size_t numInputs = 0;
    size_t numIntermediates = 0;
    std::vector<uint64_t> argStart, argLength;
    std::vector<uint64_t> outStart, outLength;
    std::vector<uint32_t> index;
    std::vector<double> consts;
    std::vector<uint32_t> outpos;
    std::vector<ProgramKernel> kernels;  // in dependency order
    int numWaves = 0;

    size_t numValues() const { return numInputs + numIntermediates; }
    size_t numResults() const {
        size_t r = 0;
        for (size_t k = 0; k < outStart.size(); ++k) r = std::max<size_t>(r, outStart[k] + outLength[k]);
        return r;
    }
};

struct ProgramInputs {
    std::unordered_map<int, size_t> inputOffsets;  // variable group -> first value
    std::unordered_map<int, size_t> inputLengths;  // variable group -> number of values
    std::vector<size_t> inputSizes;
    size_t inputDataSize = 0;
    std::vector<size_t> outputSizes;
};

struct BuildOptions {
    unsigned char decompositionThreshold = 6;
    bool simplifyTemplates = true;
};

// outputs: ASSIGN(VAR(i, -10 - container), expr) for every output value
Program buildProgram(const std::vector<Symbolic>& outputs, const ProgramInputs& io, const BuildOptions& opt);

struct LaunchDesc {
    // grid value asking the runtime for as many workgroups as the device runs at once
    static constexpr uint32_t GRID_FILL = 0xffffffffu;
    std::string name;
    uint32_t grid = 0, block = 0;
    uint32_t shared = 0;  // dynamic LDS bytes
};

struct EmitOptions {
    int threadsPerBlock = 256;
};

// HIP source with one extern "C" kernel per wave, plus the launch list.
std::string emitHip(const Program& prog, const EmitOptions& opt, std::vector<LaunchDesc>& launches);

// Host C++ source of the same program (OpenMP over instances, one parallel region per wave),
// entry point symlib_cpu_run (see symcpu.h).
std::string emitCpu(const Program& prog);

}  // namespace Sym
