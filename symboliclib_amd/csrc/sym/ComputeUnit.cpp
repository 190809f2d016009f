#include "ComputeUnit.hpp"

#include <sstream>
#include <string>

#include "Tiling.hpp"

namespace Sym {

ComputeUnitBase::ComputeUnitBase(const Device& d) : device(d), cpuDevice(!d.hipDevice) {
    if (device.cudaDevice) throw std::invalid_argument("ComputeUnit: this build has no CUDA device (use UseHIP())");
    if (device.singlePrecision) throw std::invalid_argument("ComputeUnit: this build evaluates in FP64 only");
}

ComputeUnitBase::~ComputeUnitBase() { close(); }

void ComputeUnitBase::close() {
    if (rt) {
        symhip_destroy(rt);
        rt = nullptr;
    }
    if (cpu) {
        symcpu_destroy(cpu);
        cpu = nullptr;
    }
}

void ComputeUnitBase::check(int rc, const char* what) const {
    if (rc != SYMHIP_SUCCESS)
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " +
                                 (cpuDevice ? symcpu_last_error() : symhip_last_error()));
}

// Inputs are containers of variables {0..len-1, g} of one group g; every other container is
// an output whose k-th value becomes ASSIGN(VAR(k, -10 - container), expr)
// (src/ComputeUnit.cpp:171-228).
void ComputeUnitBase::addExpressions(const Symbolic* expr, size_t len, int& id) {
    if (!len) return;
    bool isInput = true;
    for (size_t i = 0; i < len; ++i)
        if (expr[i].op() != VAR) {
            isInput = false;
            break;
        }

    if (isInput) {
        auto v0 = expr[0].variable();
        for (size_t i = 0; i < len; ++i) {
            if (expr[i].variable()[0] != (int)i)
                throw std::invalid_argument("ComputeUnit: input variables must be numbered 0..len-1 in container order");
            if (expr[i].variable()[1] != v0[1])
                throw std::invalid_argument("ComputeUnit: all variables of an input container must share one group");
        }
        if (io.inputOffsets.count(v0[1])) throw std::invalid_argument("ComputeUnit: two input containers use the same group");
        io.inputSizes.push_back(len);
        io.inputOffsets[v0[1]] = io.inputDataSize;
        io.inputLengths[v0[1]] = len;
        io.inputDataSize += len;
    } else {
        for (size_t i = 0; i < len; ++i) {
            if (expr[i].op() == NOOP) throw std::invalid_argument("ComputeUnit: output container holds an empty expression");
            outputExpressions.push_back(Symbolic(ASSIGN, Symbolic((int)i, -10 - id), expr[i]));
        }
        io.outputSizes.push_back(len);
        ++id;
    }
}

void ComputeUnitBase::init() {
    if (io.outputSizes.empty()) throw std::invalid_argument("ComputeUnit: no output expressions");
    if (device.fused) throw std::invalid_argument("ComputeUnit: UseFused applies to units built from a HessianAssembly");
    BuildOptions opt;
    opt.decompositionThreshold = (unsigned char)std::min(255u, device.decompositionThreshold);
    opt.simplifyTemplates = device.simplifyTemplates;
    program = buildProgram(outputExpressions, io, opt);
    outputExpressions.clear();  // the program holds everything needed from here on
    outputExpressions.shrink_to_fit();
    finalize();
}

void ComputeUnitBase::initFromAssembly(const HessianAssembly& assembly) {
    BuildOptions opt;
    opt.simplifyTemplates = device.simplifyTemplates;
    if (device.fused && cpuDevice) throw std::invalid_argument("ComputeUnit: the fused form runs on the UseHIP device only");
    if (device.fused) {
        FusedOptions fo;
        fo.ldsBytes = device.fusedLdsBytes;
        fo.threadsPerBlock = device.threadsPerBlock();
        if (device.schedule < 0 || device.schedule > (int)FusedSchedule::Strided) throw std::invalid_argument("ComputeUnit: unknown schedule");
        fo.schedule = (FusedSchedule)device.schedule;
        if (device.outputGroups < 1 || device.outputGroups > 64) throw std::invalid_argument("ComputeUnit: output groups must be 1..64");
        if (device.wavesPerEU < 0 || device.wavesPerEU > 8) throw std::invalid_argument("ComputeUnit: waves per EU must be 0..8");
        fo.outputGroups = device.outputGroups;
        fo.wavesPerEU = device.wavesPerEU;
        fo.profile = device.profileTiles;
        fo.split = device.fusedSplit;
        fo.ldsDma = device.ldsDma;
        fo.blockSums = device.blockSums;
        fusedProgram = assembly.buildFused(opt, fo, io, assemblyPatterns);
        fusedOptions = fo;
        code = emitHipFused(fusedProgram, fo, launches, &fusedStats);
        fusedForm = true;
        tiledForm = false;
        tilingNote = fusedProgram.summary();
        program = Program();
        program.numInputs = fusedProgram.numInputs;
        for (size_t k = 0; k < fusedProgram.outStart.size(); ++k) {
            program.outStart.push_back(fusedProgram.outStart[k]);
            program.outLength.push_back(fusedProgram.outLength[k]);
        }
        return;
    }
    program = assembly.build(opt, io, assemblyPatterns);
    finalize();
}

void ComputeUnitBase::initFromProducts(const ProductSum& ps) {
    if (cpuDevice) throw std::invalid_argument("ComputeUnit: product tiles run on the UseHIP device only");
    FusedOptions fo;
    fo.ldsBytes = device.fusedLdsBytes;
    fo.threadsPerBlock = device.threadsPerBlock();
    if (device.schedule < 0 || device.schedule > (int)FusedSchedule::Strided) throw std::invalid_argument("ComputeUnit: unknown schedule");
    fo.schedule = (FusedSchedule)device.schedule;
    fo.wavesPerEU = device.wavesPerEU;
    fo.profile = device.profileTiles;
    fo.ldsDma = device.ldsDma;
    std::vector<int> outer, inner;
    fusedProgram = buildProductTiles(ps.n, ps.colptr, ps.rowidx, fo, outer, inner);
    fusedOptions = fo;
    code = emitHipFused(fusedProgram, fo, launches, &fusedStats);
    fusedForm = true;
    tiledForm = false;
    tilingNote = fusedProgram.summary();
    io = ProgramInputs();
    io.inputSizes.push_back((size_t)ps.colptr[ps.n]);
    io.inputOffsets[0] = 0;
    io.inputLengths[0] = (size_t)ps.colptr[ps.n];
    io.inputDataSize = (size_t)ps.colptr[ps.n];
    io.outputSizes.push_back(fusedProgram.numResults);
    program = Program();
    program.numInputs = fusedProgram.numInputs;
    program.outStart.push_back(0);
    program.outLength.push_back(fusedProgram.numResults);
    SparseMatrix<double> pat(ps.n, ps.n);
    pat.setStructure(std::move(outer), std::move(inner));
    assemblyPatterns.clear();
    assemblyPatterns.push_back(std::move(pat));
}

void ComputeUnitBase::finalize() {
    EmitOptions eo;
    eo.threadsPerBlock = device.threadsPerBlock();
    tiledForm = false;
    if (cpuDevice) {
        if (device.tiled) throw std::invalid_argument("ComputeUnit: the tiled form runs on the UseHIP device only");
        code = emitCpu(program);
        launches.clear();
        return;
    }
    if (device.tiled) {
        TileOptions to;
        to.ldsBytes = device.tileLdsBytes;
        to.threadsPerBlock = eo.threadsPerBlock;
        TiledProgram tp = buildTiles(program, to);
        if (tp.valid) {
            code = emitHipTiled(program, tp, eo, launches);
            tilePool = std::move(tp.pool);
            tiledForm = true;
            std::ostringstream note;
            note << "tiles=" << tp.numTiles << " lds=" << tp.ldsBytes << " recompute="
                 << (tp.uniqueIntermediates ? (double)tp.tiledIntermediates / tp.uniqueIntermediates : 1.0);
            tilingNote = note.str();
        } else {
            tilingNote = "untiled: " + tp.reason;
        }
    }
    if (!tiledForm) code = emitHip(program, eo, launches);
}

void ComputeUnitBase::compileProgram() {
    close();
    std::vector<uint64_t> argStart(io.inputSizes.size()), argLength(io.inputSizes.size());
    {
        size_t acc = 0;
        for (size_t i = 0; i < io.inputSizes.size(); ++i) {
            argStart[i] = acc;
            argLength[i] = io.inputSizes[i];
            acc += io.inputSizes[i];
        }
    }
    std::vector<symhip_launch> ls;
    for (auto& L : launches) ls.push_back({L.name.c_str(), L.grid, L.block, L.shared});

    symhip_program_desc d{};
    d.num_values = program.numValues();
    d.num_args = (uint32_t)argStart.size();
    d.arg_start = argStart.data();
    d.arg_length = argLength.data();
    d.num_outputs = (uint32_t)program.outStart.size();
    d.out_start = program.outStart.data();
    d.out_length = program.outLength.data();
    d.num_index = program.index.size();
    d.index_data = program.index.data();
    d.num_const = program.consts.size();
    d.const_data = program.consts.data();
    d.num_outpos = tiledForm ? 0 : program.outpos.size();
    d.outpos_data = program.outpos.data();
    d.num_tiles_data = tilePool.size();
    d.tiles_data = tilePool.data();
    if (tiledForm) {
        // tiles read inputs, constants and their own tables only
        d.num_values = program.numInputs;
        d.num_index = 0;
    }
    if (fusedForm) {
        // element images of the split form live after the inputs (allocated for both forms so a
        // retune can switch between them)
        d.num_values = fusedProgram.numInputs + fusedProgram.imageDoubles;
        d.num_index = 0;
        d.num_outpos = 0;
        d.num_const = fusedProgram.consts.size();
        d.const_data = fusedProgram.consts.data();
        d.num_tiles_data = fusedProgram.pool.size();
        d.tiles_data = fusedProgram.pool.data();
    }
    d.num_work_words = fusedForm && fusedOptions.profile ? tileProfileWords(fusedProgram.numTiles) : WORK_COUNTER_WORDS;
    d.num_launches = (uint32_t)ls.size();
    d.launches = ls.data();

    if (cpuDevice) {
        symcpu_t hc = nullptr;
        check(symcpu_create(&hc, &d), "symcpu_create");
        const int rc = symcpu_compile(hc, code.c_str(), nullptr);
        if (rc) {
            const std::string msg = symcpu_last_error();
            symcpu_destroy(hc);
            throw std::runtime_error("symcpu_compile failed (" + std::to_string(rc) + "): " + msg);
        }
        cpu = hc;
        return;
    }
    symhip_t h = nullptr;
    check(symhip_create(&h, &d), "symhip_create");
    int rc = symhip_compile(h, code.c_str(), nullptr);
    if (rc) {
        const std::string msg = symhip_last_error();
        symhip_destroy(h);
        throw std::runtime_error("symhip_compile failed (" + std::to_string(rc) + "): " + msg);
    }
    rt = h;
}

void ComputeUnitBase::setArgImpl(int i, const double* d, size_t n) {
    if (!compiled()) throw std::runtime_error("ComputeUnit: execute() before compile()");
    if (i < 0 || (size_t)i >= io.inputSizes.size()) throw std::invalid_argument("ComputeUnit: too many input containers");
    if (n != io.inputSizes[i]) throw std::invalid_argument("ComputeUnit: input container " + std::to_string(i) + " has the wrong size");
    if (cpu) check(symcpu_set_arg(cpu, i, d, n), "symcpu_set_arg");
    else check(symhip_set_arg(rt, i, d, n), "symhip_set_arg");
}

void ComputeUnitBase::runImpl(void* stream) {
    if (!compiled()) throw std::runtime_error("ComputeUnit: run before compile()");
    if (cpu) check(symcpu_run(cpu, (int)std::max(1u, device.numThreads)), "symcpu_run");
    else check(symhip_run(rt, stream), "symhip_run");
}

void ComputeUnitBase::retuneFused(const FusedOptions& fo) {
    if (!fusedForm) throw std::invalid_argument("ComputeUnit: retune applies to the fused form");
    if (fo.ldsBytes != fusedOptions.ldsBytes || fo.maxOutputs != fusedOptions.maxOutputs)
        throw std::invalid_argument("ComputeUnit: the tile LDS budget cannot change without rebuilding the tiles");
    std::vector<LaunchDesc> ls;
    FusedCodeStats st;
    std::string src = emitHipFused(fusedProgram, fo, ls, &st);
    // the launch order (and the element form's header table after it) changes only between
    // order classes: build order (static, queue), largest first, interleaved (and strided)
    auto orderClass = [](FusedSchedule s) {
        return s == FusedSchedule::QueueLargest ? 1 : (s == FusedSchedule::Interleaved || s == FusedSchedule::Strided) ? 2 : 0;
    };
    const bool reorder = orderClass(fo.schedule) != orderClass(fusedOptions.schedule);
    if (reorder) orderTiles(fusedProgram, fo.schedule);
    if (rt) {
        if (reorder) {
            const size_t words = fusedProgram.numTiles + (fusedProgram.kind == 0 ? fusedProgram.numTiles * fusedProgram.headerStride : 0);
            check(symhip_update_tiles(rt, 1, words, fusedProgram.pool.data() + 1), "symhip_update_tiles");
        }
        std::vector<symhip_launch> L;
        for (auto& l : ls) L.push_back({l.name.c_str(), l.grid, l.block, l.shared});
        check(symhip_set_launches(rt, (uint32_t)L.size(), L.data()), "symhip_set_launches");
        check(symhip_set_work_words(rt, fo.profile ? tileProfileWords(fusedProgram.numTiles) : WORK_COUNTER_WORDS),
              "symhip_set_work_words");
        check(symhip_compile(rt, src.c_str(), nullptr), "symhip_compile");
    }
    code = std::move(src);
    launches = std::move(ls);
    fusedOptions = fo;
    fusedStats = st;
}

std::string ComputeUnitBase::fusedSource(const FusedOptions& fo) const {
    if (!fusedForm) throw std::invalid_argument("ComputeUnit: not a fused-form unit");
    std::vector<LaunchDesc> ls;
    return emitHipFused(fusedProgram, fo, ls);
}

std::vector<uint64_t> ComputeUnitBase::tileProfile() const {
    if (!rt || !fusedForm || !fusedOptions.profile) throw std::runtime_error("ComputeUnit: no profiled fused launch");
    std::vector<uint64_t> stamps(4 * fusedProgram.numTiles);
    check(symhip_read_work(rt, WORK_COUNTER_WORDS, 8 * fusedProgram.numTiles, reinterpret_cast<uint32_t*>(stamps.data())),
          "symhip_read_work");
    return stamps;
}

void ComputeUnitBase::runGraph(void* stream) {
    if (!rt) throw std::runtime_error("ComputeUnit: graph replay needs a compiled UseHIP unit");
    check(symhip_run_graph(rt, stream), "symhip_run_graph");
}

void ComputeUnitBase::getResultImpl(int k, double* d, size_t n) {
    if (!compiled()) throw std::runtime_error("ComputeUnit: getResults() before compile()");
    if (k < 0 || (size_t)k >= io.outputSizes.size()) throw std::invalid_argument("ComputeUnit: too many output containers");
    if (n != io.outputSizes[k]) throw std::invalid_argument("ComputeUnit: output container " + std::to_string(k) + " has the wrong size");
    if (cpu) check(symcpu_get_result(cpu, k, d, n), "symcpu_get_result");
    else check(symhip_get_result(rt, k, d, n), "symhip_get_result");
}

}  // namespace Sym
