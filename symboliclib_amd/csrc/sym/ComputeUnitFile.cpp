// ComputeUnitBase::save / initFromFile: the built program in one file (format in Persist.hpp).
#include <stdexcept>

#include "ComputeUnit.hpp"
#include "Persist.hpp"

namespace Sym {

namespace {

template <class K, class V>
std::vector<std::pair<K, V>> pairsOf(const std::unordered_map<K, V>& m) {
    return std::vector<std::pair<K, V>>(m.begin(), m.end());
}

void writeDevice(BinWriter& w, const Device& d) {
    w.pod<uint32_t>(d.vectorWidth);
    w.pod<uint32_t>(d.numThreads);
    w.pod<uint32_t>(d.decompositionThreshold);
    w.pod<uint8_t>(d.hipDevice);
    w.pod<uint8_t>(d.simplifyTemplates);
    w.pod<uint8_t>(d.tiled);
    w.pod<uint32_t>(d.tileLdsBytes);
    w.pod<uint8_t>(d.fused);
    w.pod<uint32_t>(d.fusedLdsBytes);
    w.pod<int32_t>(d.schedule);
    w.pod<int32_t>(d.outputGroups);
    w.pod<int32_t>(d.wavesPerEU);
    w.pod<uint8_t>(d.profileTiles);
    w.pod<uint8_t>(d.fusedSplit);
    w.pod<uint8_t>(d.ldsDma);
    w.pod<uint8_t>(d.blockSums);
}

void readDevice(BinReader& r, Device& d) {
    d.vectorWidth = r.pod<uint32_t>();
    d.numThreads = r.pod<uint32_t>();
    d.decompositionThreshold = r.pod<uint32_t>();
    const bool hip = r.pod<uint8_t>() != 0;
    if (hip != d.hipDevice)
        r.fail(std::string("saved for the ") + (hip ? "UseHIP" : "CPU") + " device, loaded for the " +
               (d.hipDevice ? "UseHIP" : "CPU") + " device");
    d.simplifyTemplates = r.pod<uint8_t>() != 0;
    d.tiled = r.pod<uint8_t>() != 0;
    d.tileLdsBytes = r.pod<uint32_t>();
    d.fused = r.pod<uint8_t>() != 0;
    d.fusedLdsBytes = r.pod<uint32_t>();
    d.schedule = r.pod<int32_t>();
    d.outputGroups = r.pod<int32_t>();
    d.wavesPerEU = r.pod<int32_t>();
    d.profileTiles = r.pod<uint8_t>() != 0;
    d.fusedSplit = r.pod<uint8_t>() != 0;
    d.ldsDma = r.pod<uint8_t>() != 0;
    d.blockSums = r.pod<uint8_t>() != 0;
}

void writeSizes(BinWriter& w, const std::vector<size_t>& v) {
    w.vec(std::vector<uint64_t>(v.begin(), v.end()));
}

std::vector<size_t> readSizes(BinReader& r) {
    const auto v = r.vec<uint64_t>();
    return std::vector<size_t>(v.begin(), v.end());
}

// per-kernel scalars of a ProgramKernel, in file order
struct KernelRecord {
    uint64_t n, indexOffset, constOffset, outposOffset, interBase, numVarSlots, numConstSlots, numOutSlots;
    int32_t wave, group, userOutputs, reserved;  // no padding: the file bytes are all defined
};
static_assert(sizeof(KernelRecord) == 80, "KernelRecord has padding");

}  // namespace

void ComputeUnitBase::save(const std::string& path) const {
    BinWriter w(path);
    w.raw(PROGRAM_FILE_MAGIC, 8);
    w.pod<uint32_t>(PROGRAM_FILE_VERSION);
    writeDevice(w, device);
    w.pod<uint8_t>(tiledForm);
    w.pod<uint8_t>(fusedForm);
    w.pod<int32_t>(numOutputContainers);

    // input and output layout
    writeSizes(w, io.inputSizes);
    writeSizes(w, io.outputSizes);
    w.pod<uint64_t>(io.inputDataSize);
    std::vector<int64_t> offs, lens;
    for (const auto& kv : pairsOf(io.inputOffsets)) {
        offs.push_back(kv.first);
        offs.push_back((int64_t)kv.second);
    }
    for (const auto& kv : pairsOf(io.inputLengths)) {
        lens.push_back(kv.first);
        lens.push_back((int64_t)kv.second);
    }
    w.vec(offs);
    w.vec(lens);

    // program tables and kernels
    w.pod<uint64_t>(program.numInputs);
    w.pod<uint64_t>(program.numIntermediates);
    w.vec(program.argStart);
    w.vec(program.argLength);
    w.vec(program.outStart);
    w.vec(program.outLength);
    w.vec(program.index);
    w.vec(program.consts);
    w.vec(program.outpos);
    w.pod<int32_t>(program.numWaves);
    std::vector<KernelRecord> krec;
    std::vector<Symbolic> codes;
    std::vector<uint64_t> accessCount;
    std::vector<ColumnRef> access;
    for (const ProgramKernel& k : program.kernels) {
        krec.push_back({k.n, k.indexOffset, k.constOffset, k.outposOffset, k.interBase, k.numVarSlots, k.numConstSlots,
                        k.numOutSlots, k.wave, k.group, (int32_t)k.userOutputs, 0});
        codes.push_back(k.code);
        accessCount.push_back(k.access.size());
        access.insert(access.end(), k.access.begin(), k.access.end());
    }
    w.vec(krec);
    w.vec(accessCount);
    w.vec(access);

    // emitted source, launches, tile pool of the tiled form
    w.str(code);
    w.pod<uint64_t>(launches.size());
    for (const LaunchDesc& l : launches) {
        w.str(l.name);
        w.pod<uint32_t>(l.grid);
        w.pod<uint32_t>(l.block);
        w.pod<uint32_t>(l.shared);
    }
    w.vec(tilePool);
    w.str(tilingNote);

    // fused program
    const FusedProgram& F = fusedProgram;
    w.pod<int32_t>(F.kind);
    w.pod<uint64_t>(F.templates.size());
    for (const FusedTemplate& t : F.templates) {
        codes.push_back(t.code);
        w.pod<int32_t>(t.numNodes);
        w.vec(t.argNode);
        w.vec(t.argDof);
        w.pod<int32_t>(t.numParams);
        w.pod<int32_t>(t.numOutputs);
        w.pod<int32_t>(t.valueStride);
    }
    w.vec(F.pool);
    w.vec(F.consts);
    w.vec(F.tileHeader);
    w.vec(F.tileInstance);
    w.vec(F.tileIndex);
    w.vec(F.tileCost);
    writeSizes(w, F.outStart);
    writeSizes(w, F.outLength);
    for (size_t v : {F.numInputs, F.numResults, F.numTiles, F.ldsBytes, F.elementLdsBytes, F.stagedLdsBytes,
                     F.imageDoubles, F.headerTable, F.headerStride, F.headerNodes, F.elementEvaluations, F.numElements,
                     F.numContributions})
        w.pod<uint64_t>(v);
    w.pod<int32_t>(F.stagedDofs);
    w.pod<int32_t>(F.maxPairs);
    w.pod<int32_t>((int)F.sums);
    w.pod<int32_t>(F.blockDofs);
    w.pod<uint64_t>(F.maxBlocks);
    const FusedOptions& O = fusedOptions;
    for (size_t v : {O.ldsBytes, O.maxOutputs}) w.pod<uint64_t>(v);
    for (int v : {O.threadsPerBlock, (int)O.schedule, O.outputGroups, O.wavesPerEU}) w.pod<int32_t>(v);
    for (bool v : {O.profile, O.shareReciprocals, O.foldCoefficients, O.split, O.ldsDma, O.blockSums}) w.pod<uint8_t>(v);
    const FusedCodeStats& S = fusedStats;
    for (size_t v : {S.adds, S.muls, S.divs, S.sqrts, S.statements}) w.pod<uint64_t>(v);
    w.pod<int32_t>(S.maxLive);

    // output patterns, then every template expression in one node table
    w.pod<uint64_t>(assemblyPatterns.size());
    for (const auto& P : assemblyPatterns) {
        w.pod<int32_t>(P.rows());
        w.pod<int32_t>(P.cols());
        w.vec(P.outer());
        w.vec(P.inner());
    }
    w.symbolics(codes);
    w.raw(PROGRAM_FILE_MAGIC, 8);
    w.finish();
}

void ComputeUnitBase::initFromFile(const std::string& path) {
    BinReader r(path);
    char magic[8];
    r.raw(magic, 8);
    if (std::string(magic, 8) != std::string(PROGRAM_FILE_MAGIC, 8)) r.fail("not a program file");
    if (r.pod<uint32_t>() != PROGRAM_FILE_VERSION) r.fail("unsupported version");
    readDevice(r, device);
    tiledForm = r.pod<uint8_t>() != 0;
    fusedForm = r.pod<uint8_t>() != 0;
    numOutputContainers = r.pod<int32_t>();

    io = ProgramInputs();
    io.inputSizes = readSizes(r);
    io.outputSizes = readSizes(r);
    io.inputDataSize = r.pod<uint64_t>();
    const auto offs = r.vec<int64_t>(), lens = r.vec<int64_t>();
    if (offs.size() % 2 || lens.size() % 2) r.fail("bad input layout");
    for (size_t i = 0; i < offs.size(); i += 2) io.inputOffsets[(int)offs[i]] = (size_t)offs[i + 1];
    for (size_t i = 0; i < lens.size(); i += 2) io.inputLengths[(int)lens[i]] = (size_t)lens[i + 1];

    program = Program();
    program.numInputs = r.pod<uint64_t>();
    program.numIntermediates = r.pod<uint64_t>();
    program.argStart = r.vec<uint64_t>();
    program.argLength = r.vec<uint64_t>();
    program.outStart = r.vec<uint64_t>();
    program.outLength = r.vec<uint64_t>();
    program.index = r.vec<uint32_t>();
    program.consts = r.vec<double>();
    program.outpos = r.vec<uint32_t>();
    program.numWaves = r.pod<int32_t>();
    const auto krec = r.vec<KernelRecord>();
    const auto accessCount = r.vec<uint64_t>();
    const auto access = r.vec<ColumnRef>();
    if (accessCount.size() != krec.size()) r.fail("bad kernel table");
    size_t a = 0;
    for (size_t i = 0; i < krec.size(); ++i) {
        const KernelRecord& k = krec[i];
        ProgramKernel P;
        P.n = k.n;
        P.wave = k.wave;
        P.userOutputs = k.userOutputs != 0;
        P.indexOffset = k.indexOffset;
        P.constOffset = k.constOffset;
        P.outposOffset = k.outposOffset;
        P.interBase = k.interBase;
        P.numVarSlots = k.numVarSlots;
        P.numConstSlots = k.numConstSlots;
        P.numOutSlots = k.numOutSlots;
        P.group = k.group;
        if (accessCount[i] > access.size() - a) r.fail("bad kernel access table");
        P.access.assign(access.begin() + (long)a, access.begin() + (long)(a + accessCount[i]));
        a += accessCount[i];
        program.kernels.push_back(std::move(P));
    }

    code = r.str();
    const uint64_t nl = r.pod<uint64_t>();
    launches.clear();
    for (uint64_t i = 0; i < nl; ++i) {
        LaunchDesc l;
        l.name = r.str();
        l.grid = r.pod<uint32_t>();
        l.block = r.pod<uint32_t>();
        l.shared = r.pod<uint32_t>();
        launches.push_back(std::move(l));
    }
    tilePool = r.vec<uint32_t>();
    tilingNote = r.str();

    FusedProgram& F = fusedProgram;
    F = FusedProgram();
    F.kind = r.pod<int32_t>();
    const uint64_t nt = r.pod<uint64_t>();
    for (uint64_t i = 0; i < nt; ++i) {
        FusedTemplate t;
        t.numNodes = r.pod<int32_t>();
        t.argNode = r.vec<int>();
        t.argDof = r.vec<int>();
        t.numParams = r.pod<int32_t>();
        t.numOutputs = r.pod<int32_t>();
        t.valueStride = r.pod<int32_t>();
        F.templates.push_back(std::move(t));
    }
    F.pool = r.vec<uint32_t>();
    F.consts = r.vec<double>();
    F.tileHeader = r.vec<uint32_t>();
    F.tileInstance = r.vec<uint32_t>();
    F.tileIndex = r.vec<uint32_t>();
    F.tileCost = r.vec<float>();
    F.outStart = readSizes(r);
    F.outLength = readSizes(r);
    for (size_t* v : {&F.numInputs, &F.numResults, &F.numTiles, &F.ldsBytes, &F.elementLdsBytes, &F.stagedLdsBytes,
                      &F.imageDoubles, &F.headerTable, &F.headerStride, &F.headerNodes, &F.elementEvaluations,
                      &F.numElements, &F.numContributions})
        *v = r.pod<uint64_t>();
    F.stagedDofs = r.pod<int32_t>();
    F.maxPairs = r.pod<int32_t>();
    const int sums = r.pod<int32_t>();
    if (sums != (int)FusedSums::Outputs && sums != (int)FusedSums::Blocks) r.fail("unknown sum form");
    F.sums = (FusedSums)sums;
    F.blockDofs = r.pod<int32_t>();
    F.maxBlocks = r.pod<uint64_t>();
    FusedOptions& O = fusedOptions;
    O.ldsBytes = r.pod<uint64_t>();
    O.maxOutputs = r.pod<uint64_t>();
    O.threadsPerBlock = r.pod<int32_t>();
    const int sched = r.pod<int32_t>();
    if (sched < 0 || sched > (int)FusedSchedule::Strided) r.fail("unknown schedule");
    O.schedule = (FusedSchedule)sched;
    O.outputGroups = r.pod<int32_t>();
    O.wavesPerEU = r.pod<int32_t>();
    for (bool* v : {&O.profile, &O.shareReciprocals, &O.foldCoefficients, &O.split, &O.ldsDma, &O.blockSums})
        *v = r.pod<uint8_t>() != 0;
    FusedCodeStats& S = fusedStats;
    for (size_t* v : {&S.adds, &S.muls, &S.divs, &S.sqrts, &S.statements}) *v = r.pod<uint64_t>();
    S.maxLive = r.pod<int32_t>();

    const uint64_t np = r.pod<uint64_t>();
    assemblyPatterns.clear();
    for (uint64_t i = 0; i < np; ++i) {
        const int rows = r.pod<int32_t>(), cols = r.pod<int32_t>();
        auto outer = r.vec<int>();
        auto inner = r.vec<int>();
        if (rows < 0 || cols < 0 || outer.size() != (size_t)cols + 1 || outer.front() != 0 ||
            (size_t)outer.back() != inner.size())
            r.fail("bad output pattern");
        SparseMatrix<double> P(rows, cols);
        P.setStructure(std::move(outer), std::move(inner));
        assemblyPatterns.push_back(std::move(P));
    }
    const std::vector<Symbolic> codes = r.symbolics();
    if (codes.size() != program.kernels.size() + F.templates.size()) r.fail("bad template table");
    for (size_t i = 0; i < program.kernels.size(); ++i) program.kernels[i].code = codes[i];
    for (size_t i = 0; i < F.templates.size(); ++i) F.templates[i].code = codes[program.kernels.size() + i];
    char tail[8];
    r.raw(tail, 8);
    if (std::string(tail, 8) != std::string(PROGRAM_FILE_MAGIC, 8)) r.fail("bad end marker");
    // every index of a tile pool against its buffers, as after a build
    if (fusedForm) {
        try {
            validateFused(F);
        } catch (const std::logic_error& e) {
            r.fail(e.what());
        }
    }
}

}  // namespace Sym
