// ComputeUnit: compile a set of symbolic output expressions over symbolic input containers
// into an MI355X program, then evaluate it for numeric inputs.
//
// Drop-in for the reference's UseHIP path (src/ComputeUnit.hpp:14-223):
//
//   Sym::ComputeUnit<double> unit(Sym::Device(Sym::UseHIP(), Sym::ThreadsPerBlock(256)), AS, BS);
//   unit.compile().execute(A).getResults(B);
//
// Constructor arguments after the Device are containers in order; containers holding only
// variables {0..len-1, group} are inputs, all others are outputs.  execute() takes the
// numeric inputs in the same order, getResults() fills the outputs in the same order.
// Underneath, the unit calls the C runtime declared in symhip.h: hipRTC compilation in
// process, device-resident tables, one launch per wave of independent kernels.  A Device
// without UseHIP() is the reference's CPU device (VecWidth, NumThreads): the same program
// emitted as host C++ with OpenMP, compiled with the reference's flags (symcpu.h); it runs the
// waves form only.
//
// Differences from the reference: errors throw (std::invalid_argument, std::runtime_error)
// instead of printing and degrading to no-ops; RealT = double only; no CUDA device.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "Assembly.hpp"
#include "ContainerSupport.hpp"
#include "Program.hpp"
#include "Symbolic.hpp"
#include "symcpu.h"
#include "symhip.h"

namespace Sym {

template <class T, class Tag>
struct NamedType {
    T x;
    explicit NamedType(const T& x_) : x(x_) {}
};

using ThreadsPerBlock = NamedType<int, struct ThreadsPerBlockTag>;
using NumThreads = NamedType<int, struct NumThreadsTag>;
using VecWidth = NamedType<int, struct VecWidthTag>;
using DecompositionThreshold = NamedType<int, struct DecompositionThresholdTag>;

struct UseHIP {};
struct UseCuda {};
struct UseSinglePrecision {};
// Skip algebraic simplification of kernel templates (faster compile of very large templates).
struct NoSimplify {};
// Execution form: one launch per wave of kernels with intermediates in HBM (default), or
// LDS-staged work tiles (UseTiles; used when every output's closure fits a tile).
struct UseWaves {};
struct UseTiles {};
using TileLDSBytes = NamedType<int, struct TileLDSBytesTag>;
// Fused assembly form of a unit built from a HessianAssembly: tiles of owner nodes evaluate
// their elements into LDS and write their outputs directly (one launch).  Schedule picks the
// tile order and whether workgroups are persistent (FusedSchedule values).
struct UseFused {};
using Schedule = NamedType<int, struct ScheduleTag>;
// Fused form: work items per element (its outputs split between them) and the minimum
// waves per SIMD the kernel is compiled for (0 = compiler's choice).
using OutputGroups = NamedType<int, struct OutputGroupsTag>;
using WavesPerEU = NamedType<int, struct WavesPerEUTag>;
// Fused form in two launches: element values of every tile to HBM, then the sums.
struct UseSplit {};
// Fused form: stage tiles through registers instead of direct global-to-LDS loads.
struct UseRegisterStaging {};
// Fused form: one contribution list per output even when the Hessian is made of node blocks.
struct UseOutputSums {};

class Device {
public:
    unsigned int vectorWidth = 1;
    unsigned int numThreads = 1;
    unsigned int decompositionThreshold = 6;
    bool singlePrecision = false;
    bool cudaDevice = false;
    bool hipDevice = false;
    bool simplifyTemplates = true;
    bool tiled = false;
    unsigned int tileLdsBytes = 64 * 1024;
    bool fused = false;
    unsigned int fusedLdsBytes = 53760;
    int schedule = (int)FusedSchedule::Queue;
    int outputGroups = 5;
    int wavesPerEU = 0;
    bool profileTiles = false;
    bool fusedSplit = false;
    bool ldsDma = true;
    bool blockSums = true;

private:
    void set(const DecompositionThreshold& nt) { decompositionThreshold = (unsigned)nt.x; }
    void set(const ThreadsPerBlock& nt) { numThreads = (unsigned)nt.x; }
    void set(const VecWidth& nt) { vectorWidth = (unsigned)nt.x; }
    void set(const NumThreads& nt) { numThreads = (unsigned)nt.x; }
    void set(const UseSinglePrecision&) { singlePrecision = true; }
    void set(const NoSimplify&) { simplifyTemplates = false; }
    void set(const UseWaves&) { tiled = false; }
    void set(const UseTiles&) { tiled = true; }
    void set(const TileLDSBytes& b) {
        tileLdsBytes = (unsigned)b.x;
        fusedLdsBytes = (unsigned)b.x;
    }
    void set(const UseFused&) { fused = true; }
    void set(const Schedule& s) { schedule = s.x; }
    void set(const OutputGroups& g) { outputGroups = g.x; }
    void set(const WavesPerEU& w) { wavesPerEU = w.x; }
    void set(const UseSplit&) { fusedSplit = true; }
    void set(const UseRegisterStaging&) { ldsDma = false; }
    void set(const UseOutputSums&) { blockSums = false; }
    void set(const UseHIP&) {
        hipDevice = true;
        cudaDevice = false;
        vectorWidth = 0;
    }
    void set(const UseCuda&) {
        cudaDevice = true;
        hipDevice = false;
        vectorWidth = 0;
    }
    template <class T>
    void set(const T&) {
        throw std::invalid_argument("Device: parameter not understood");
    }
    void init() {}
    template <class T, class... Args>
    void init(const T& t, const Args&... args) {
        set(t);
        init(args...);
    }

public:
    template <class... Args>
    Device(const Args&... args) {
        init(args...);
    }

    int threadsPerBlock() const { return numThreads <= 1 ? 256 : (int)numThreads; }
};

// B = A^T A + A for a sparse A given by its CSC pattern (values are the unit's input), built
// directly in the fused product-sum form (no per-nonzero expressions).
struct ProductSum {
    int n;
    const int* colptr;
    const int* rowidx;
};

// A program saved with ComputeUnitBase::save, to construct a unit from without rebuilding it.
struct SavedProgram {
    std::string path;
};

// Everything of ComputeUnit that does not depend on the scalar type.
class ComputeUnitBase {
protected:
    Device device;
    std::vector<Symbolic> outputExpressions;
    ProgramInputs io;
    Program program;
    std::string code;
    std::vector<LaunchDesc> launches;
    std::vector<uint32_t> tilePool;
    bool tiledForm = false;
    FusedProgram fusedProgram;
    FusedOptions fusedOptions;
    FusedCodeStats fusedStats;
    bool fusedForm = false;
    std::string tilingNote;
    bool cpuDevice = false;  // a Device without UseHIP(): host code through symcpu
    symhip_t rt = nullptr;
    symcpu_t cpu = nullptr;
    int numOutputContainers = 0;

    std::vector<SparseMatrix<double>> assemblyPatterns;

    void addExpressions(const Symbolic* expr, size_t len, int& id);
    void init();
    void initFromAssembly(const HessianAssembly& assembly);
    void initFromProducts(const ProductSum& ps);
    void initFromFile(const std::string& path);
    void finalize();
    void setArgImpl(int i, const double* d, size_t n);
    void getResultImpl(int k, double* d, size_t n);
    void runImpl(void* stream);
    void check(int rc, const char* what) const;

    explicit ComputeUnitBase(const Device& d);

public:
    ComputeUnitBase(const ComputeUnitBase&) = delete;
    ComputeUnitBase& operator=(const ComputeUnitBase&) = delete;
    virtual ~ComputeUnitBase();

    void close();
    // Write the built program (tables, templates, source, launches, tile pools, patterns) to
    // `path`; a unit constructed from SavedProgram{path} is the same program.
    void save(const std::string& path) const;
    bool compiled() const { return rt != nullptr || cpu != nullptr; }
    bool onCpu() const { return cpuDevice; }

    const Program& getProgram() const { return program; }
    const std::string& getCode() const { return code; }
    const std::vector<LaunchDesc>& getLaunches() const { return launches; }
    bool isTiled() const { return tiledForm; }
    bool isFused() const { return fusedForm; }
    const FusedProgram& getFusedProgram() const { return fusedProgram; }
    const FusedOptions& getFusedOptions() const { return fusedOptions; }
    const FusedCodeStats& getFusedStats() const { return fusedStats; }
    // Fused form: new output groups, waves per EU, block size, schedule or profiling without
    // rebuilding the tiles (ldsBytes must stay); recompiles when already compiled.
    void retuneFused(const FusedOptions& fo);
    // Fused form: the source the options would emit (for compiling ahead in parallel).
    std::string fusedSource(const FusedOptions& fo) const;
    // Profiled fused launch: four wall-clock stamps per tile in launch order.
    std::vector<uint64_t> tileProfile() const;
    // output patterns of a unit built from a HessianAssembly (one per instance)
    const std::vector<SparseMatrix<double>>& outputPatterns() const { return assemblyPatterns; }
    // record the output patterns of a unit built from expressions (kept by save())
    void setOutputPatterns(std::vector<SparseMatrix<double>> p) { assemblyPatterns = std::move(p); }
    const std::string& tilingInfo() const { return tilingNote; }
    size_t tilePoolSize() const { return tilePool.size(); }
    symhip_t runtime() const { return rt; }
    size_t numInputs() const { return io.inputSizes.size(); }
    size_t inputSize(int i) const { return io.inputSizes.at(i); }
    size_t numOutputs() const { return io.outputSizes.size(); }
    size_t outputSize(int k) const { return io.outputSizes.at(k); }

    // compile (hipRTC), upload tables
    void compileProgram();
    // run on a stream without synchronising (inputs already on the device)
    void run(void* stream = nullptr) { runImpl(stream); }
    // same, replayed from a HIP graph captured on the first call
    void runGraph(void* stream = nullptr);
};

template <class RealT>
class ComputeUnit : public ComputeUnitBase {
    static_assert(std::is_same<RealT, double>::value, "this build evaluates programs in FP64 only (RealT = double)");

    template <class... TArgs>
    void setExpressions(int) {}

    template <class T, class... TArgs>
    void setExpressions(int id, const T& A, const TArgs&... args) {
        addExpressions(contPtr(A), contSize(A), id);
        setExpressions(id, args...);
    }

    void setArgs(int) {}
    template <class T, class... TArg>
    void setArgs(int id, const T& arg, const TArg&... rest) {
        setArgImpl(id, contPtr(arg), contSize(arg));
        setArgs(id + 1, rest...);
    }

    void getRes(int) {}
    template <class T, class... TArg>
    void getRes(int id, T& arg, TArg&... rest) {
        getResultImpl(id, contPtr(arg), contSize(arg));
        getRes(id + 1, rest...);
    }

public:
    template <class... TArg>
    ComputeUnit(Device device_, const TArg&... expressions) : ComputeUnitBase(device_) {
        int id = 0;
        setExpressions(id, expressions...);
        init();
    }

    // Unit of a directly assembled element-sum Hessian (one input and one output container per
    // assembly instance, in instance order).
    ComputeUnit(Device device_, const HessianAssembly& assembly) : ComputeUnitBase(device_) {
        initFromAssembly(assembly);
    }

    // Unit of B = A^T A + A in the fused product-sum form (input: A's values, output: B's).
    ComputeUnit(Device device_, const ProductSum& ps) : ComputeUnitBase(device_) { initFromProducts(ps); }

    // A program written by save(): no front-end, no build; compile() finds its code object in
    // the runtime's cache.  The device kind (UseHIP or the CPU device) must be the saved one;
    // the other Device settings come from the file.
    ComputeUnit(Device device_, const SavedProgram& saved) : ComputeUnitBase(device_) { initFromFile(saved.path); }

    ComputeUnit& compile() {
        compileProgram();
        return *this;
    }

    template <class... TArg>
    ComputeUnit& execute(const TArg&... inputs) {
        setArgs(0, inputs...);
        runImpl(nullptr);
        return *this;
    }

    template <class TArg>
    ComputeUnit& getResults(TArg& out, const int i) {
        getResultImpl(i, contPtr(out), contSize(out));
        return *this;
    }

    template <class... TArg>
    ComputeUnit& getResults(TArg&... outs) {
        getRes(0, outs...);
        return *this;
    }

    // pointer variants for bindings
    ComputeUnit& setInput(int i, const double* d, size_t n) {
        setArgImpl(i, d, n);
        return *this;
    }
    ComputeUnit& getOutput(int k, double* d, size_t n) {
        getResultImpl(k, d, n);
        return *this;
    }
};

}  // namespace Sym
