// C ABI over ComputeUnit for the benchmark workloads (symlib.h).
#include "symlib.h"

#include <chrono>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>

#include "../sym/ComputeUnit.hpp"
#include "../sym/SymbolicDifferentiation.hpp"
#include "../workloads/Workloads.hpp"

namespace {

thread_local std::string g_err;

using Clock = std::chrono::steady_clock;
double seconds(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    } catch (...) {
        g_err = "unknown C++ exception";
        return 1;
    }
}

Sym::Device makeDevice(const symlib_options* opt) {
    symlib_options o;
    symlib_default_options(&o);
    if (opt) o = *opt;
    // cpu_threads > 0: the reference's CPU device, Device(VecWidth(4), NumThreads(n))
    Sym::Device d = o.cpu_threads > 0 ? Sym::Device(Sym::VecWidth(4), Sym::NumThreads(o.cpu_threads))
                                      : Sym::Device(Sym::UseHIP(), Sym::ThreadsPerBlock(o.threads_per_block > 0 ? o.threads_per_block : 256));
    if (o.decomposition_threshold >= 0) d.decompositionThreshold = (unsigned)o.decomposition_threshold;
    d.simplifyTemplates = o.simplify != 0;
    d.tiled = o.tiled != 0;
    if (o.tile_lds_bytes > 0) {
        d.tileLdsBytes = (unsigned)o.tile_lds_bytes;
        d.fusedLdsBytes = (unsigned)o.tile_lds_bytes;
    }
    d.fused = o.fused != 0;
    d.schedule = o.schedule;
    if (o.output_groups > 0) d.outputGroups = o.output_groups;
    d.wavesPerEU = o.waves_per_eu;
    d.profileTiles = o.profile_tiles != 0;
    d.fusedSplit = o.split != 0;
    d.ldsDma = o.register_staging == 0;
    d.blockSums = o.output_sums == 0;
    return d;
}

}  // namespace

struct symlib_unit {
    std::unique_ptr<Sym::ComputeUnit<double>> unit;
    std::string scratch;  // fused_source result
    // output patterns (kept by the unit, so saved with it)
    const std::vector<Sym::SparseMatrix<double>>& patterns() const { return unit->outputPatterns(); }
    symlib_unit_info info{};
};

namespace {

template <class... Containers>
void finishUnit(symlib_unit* u, const Sym::Device& dev, Clock::time_point t0, Clock::time_point t1,
                const Containers&... cs) {
    u->unit.reset(new Sym::ComputeUnit<double>(dev, cs...));
    const Clock::time_point t2 = Clock::now();
    const auto& P = u->unit->getProgram();
    u->info.num_inputs = P.numInputs;
    u->info.num_intermediates = P.numIntermediates;
    u->info.num_index = P.index.size();
    u->info.num_const = P.consts.size();
    u->info.num_outpos = P.outpos.size();
    u->info.num_results = P.numResults();
    u->info.num_kernels = (uint32_t)P.kernels.size();
    u->info.num_waves = (uint32_t)P.numWaves;
    u->info.source_bytes = u->unit->getCode().size();
    u->info.tiled = u->unit->isTiled() ? 1 : 0;
    u->info.tile_pool_entries = u->unit->tilePoolSize();
    std::snprintf(u->info.tiling, sizeof(u->info.tiling), "%s", u->unit->tilingInfo().c_str());
    u->info.frontend_seconds = seconds(t0, t1);
    u->info.build_seconds = seconds(t1, t2);
}

void fillInfo(symlib_unit* u) {
    const auto& P = u->unit->getProgram();
    u->info.num_inputs = P.numInputs;
    u->info.num_intermediates = P.numIntermediates;
    u->info.num_index = P.index.size();
    u->info.num_const = P.consts.size();
    u->info.num_outpos = P.outpos.size();
    u->info.num_results = P.numResults();
    u->info.num_kernels = (uint32_t)P.kernels.size();
    u->info.num_waves = (uint32_t)P.numWaves;
    u->info.source_bytes = u->unit->getCode().size();
    u->info.tiled = u->unit->isTiled() ? 1 : 0;
    u->info.tile_pool_entries = u->unit->tilePoolSize();
    if (u->unit->isFused()) {
        const auto& F = u->unit->getFusedProgram();
        u->info.num_const = F.consts.size();
        u->info.num_results = F.numResults;
        u->info.num_kernels = (uint32_t)F.templates.size();
        u->info.num_waves = 1;
        u->info.tiled = 2;
        u->info.tile_pool_entries = F.pool.size();
    }
    std::snprintf(u->info.tiling, sizeof(u->info.tiling), "%s", u->unit->tilingInfo().c_str());
}

void finishAssembly(symlib_unit* u, const Sym::Device& dev, Clock::time_point t0, const Sym::HessianAssembly& A) {
    const Clock::time_point t1 = Clock::now();
    u->unit.reset(new Sym::ComputeUnit<double>(dev, A));
    const Clock::time_point t2 = Clock::now();
    fillInfo(u);
    u->info.frontend_seconds = seconds(t0, t1);
    u->info.build_seconds = seconds(t1, t2);
}

void finishProducts(symlib_unit* u, const Sym::Device& dev, Clock::time_point t0, const Sym::ProductSum& ps) {
    const Clock::time_point t1 = Clock::now();
    u->unit.reset(new Sym::ComputeUnit<double>(dev, ps));
    const Clock::time_point t2 = Clock::now();
    const auto& F = u->unit->getFusedProgram();
    u->info.num_inputs = F.numInputs;
    u->info.num_results = F.numResults;
    u->info.num_const = 0;
    u->info.num_kernels = 0;
    u->info.num_waves = 1;
    u->info.tiled = 2;
    u->info.tile_pool_entries = F.pool.size();
    u->info.source_bytes = u->unit->getCode().size();
    std::snprintf(u->info.tiling, sizeof(u->info.tiling), "%s", u->unit->tilingInfo().c_str());
    u->info.frontend_seconds = seconds(t0, t1);
    u->info.build_seconds = seconds(t1, t2);
}

Sym::workloads::Faces toFaces(int nf, const int* faces) {
    Sym::workloads::Faces F((size_t)nf);
    for (int i = 0; i < nf; ++i) F[i] = {faces[3 * i], faces[3 * i + 1], faces[3 * i + 2]};
    return F;
}

Sym::SparseMatrix<double> patternOf(const Sym::SparseMatrix<Sym::Symbolic>& S) {
    Sym::SparseMatrix<double> P(S.rows(), S.cols());
    P.setStructure(S.outer(), S.inner());
    return P;
}

void checkFaces(int nv, int nf, const int* faces) {
    if (nv <= 0 || nf <= 0 || !faces) throw std::invalid_argument("empty mesh");
    for (long i = 0; i < 3L * nf; ++i)
        if (faces[i] < 0 || faces[i] >= nv) throw std::invalid_argument("face index out of range");
}

}  // namespace

extern "C" {

const char* symlib_last_error(void) { return g_err.c_str(); }

void symlib_default_options(symlib_options* opt) {
    if (!opt) return;
    opt->threads_per_block = 256;
    opt->decomposition_threshold = -1;
    opt->simplify = 1;
    opt->share_templates = 1;
    opt->tiled = 0;
    opt->tile_lds_bytes = 0;
    opt->direct_assembly = 0;
    opt->fused = 0;
    opt->schedule = 1;
    opt->output_groups = 0;
    opt->waves_per_eu = 0;
    opt->profile_tiles = 0;
    opt->split = 0;
    opt->register_staging = 0;
    opt->output_sums = 0;
    opt->cpu_threads = 0;
}

int symlib_unit_ata_plus_a(symlib_unit_t* out, int n, const int* colptr, const int* rowidx, const symlib_options* opt) {
    if (!out) return 1;
    *out = nullptr;
    return guarded([&] {
        if (n <= 0 || !colptr || !rowidx) throw std::invalid_argument("empty matrix");
        for (int j = 0; j < n; ++j)
            if (colptr[j + 1] < colptr[j]) throw std::invalid_argument("column pointers must not decrease");
        std::unique_ptr<symlib_unit> u(new symlib_unit());
        const auto t0 = Clock::now();
        for (int j = 0; j < n; ++j)
            for (int p = colptr[j]; p < colptr[j + 1]; ++p)
                if (rowidx[p] < 0 || rowidx[p] >= n || (p > colptr[j] && rowidx[p] <= rowidx[p - 1]))
                    throw std::invalid_argument("row indices must be in range, sorted and unique per column");
        if (opt && opt->fused) {
            finishProducts(u.get(), makeDevice(opt), t0, Sym::ProductSum{n, colptr, rowidx});
            *out = u.release();
            return;
        }
        Sym::SparseMatrix<double> A(n, n, colptr, rowidx, nullptr);
        auto AS = Sym::makeSymbolic(A, 0);
        auto BS = Sym::workloads::ataPlusA(AS);
        const auto t1 = Clock::now();
        finishUnit(u.get(), makeDevice(opt), t0, t1, AS, BS);
        u->unit->setOutputPatterns({patternOf(BS)});
        *out = u.release();
    });
}

int symlib_unit_cotan(symlib_unit_t* out, int nv, int nf, const int* faces, const symlib_options* opt) {
    if (!out) return 1;
    *out = nullptr;
    return guarded([&] {
        checkFaces(nv, nf, faces);
        std::unique_ptr<symlib_unit> u(new symlib_unit());
        const auto t0 = Clock::now();
        if (opt && (opt->direct_assembly || opt->fused)) {
            Sym::HessianAssembly A;
            Sym::workloads::CotanShape shape;
            Sym::workloads::cotanAssembly(A, shape, nv, toFaces(nf, faces), 0);
            finishAssembly(u.get(), makeDevice(opt), t0, A);
            *out = u.release();
            return;
        }
        Sym::DenseMatrix<double> V(nv, 3);
        auto VS = Sym::makeSymbolic(V, 0);
        Sym::SparseMatrix<Sym::Symbolic> LS;
        Sym::workloads::buildCotan(VS, toFaces(nf, faces), LS);
        const auto t1 = Clock::now();
        finishUnit(u.get(), makeDevice(opt), t0, t1, VS, LS);
        u->unit->setOutputPatterns({patternOf(LS)});
        *out = u.release();
    });
}

int symlib_unit_bw_hessian(symlib_unit_t* out, int nv, int nf, const int* faces, const double* rest, double k_stretch,
                           double k_shear, double k_bend, const symlib_options* opt) {
    if (!out) return 1;
    *out = nullptr;
    return guarded([&] {
        checkFaces(nv, nf, faces);
        if (!rest) throw std::invalid_argument("rest positions missing");
        std::unique_ptr<symlib_unit> u(new symlib_unit());
        const auto t0 = Clock::now();
        Sym::DenseMatrix<double> R(nv, 3, rest);
        if (opt && (opt->direct_assembly || opt->fused)) {
            Sym::HessianAssembly A;
            Sym::workloads::FaceShapes shapes(k_stretch, k_shear);
            Sym::workloads::baraffWitkinAssembly(A, shapes, R, toFaces(nf, faces), k_bend, 0);
            finishAssembly(u.get(), makeDevice(opt), t0, A);
            *out = u.release();
            return;
        }
        Sym::DenseMatrix<double> X(nv, 3);
        auto XS = Sym::makeSymbolic(X, 0);
        const bool share = opt ? opt->share_templates != 0 : true;
        auto HS = Sym::workloads::baraffWitkinHessian(XS, R, toFaces(nf, faces), k_stretch, k_shear, k_bend, share);
        const auto t1 = Clock::now();
        finishUnit(u.get(), makeDevice(opt), t0, t1, XS, HS);
        u->unit->setOutputPatterns({patternOf(HS)});
        *out = u.release();
    });
}

int symlib_unit_bw_hessian_batch(symlib_unit_t* out, int count, const int* nv, const int* nf, const int* const* faces,
                                 const double* const* rest, double k_stretch, double k_shear, double k_bend,
                                 const symlib_options* opt) {
    if (!out) return 1;
    *out = nullptr;
    return guarded([&] {
        if (count <= 0 || !nv || !nf || !faces || !rest) throw std::invalid_argument("empty batch");
        std::unique_ptr<symlib_unit> u(new symlib_unit());
        const auto t0 = Clock::now();
        Sym::HessianAssembly A;
        Sym::workloads::FaceShapes shapes(k_stretch, k_shear);
        for (int k = 0; k < count; ++k) {
            checkFaces(nv[k], nf[k], faces[k]);
            if (!rest[k]) throw std::invalid_argument("rest positions missing");
            Sym::DenseMatrix<double> R(nv[k], 3, rest[k]);
            Sym::workloads::baraffWitkinAssembly(A, shapes, R, toFaces(nf[k], faces[k]), k_bend, k);
        }
        finishAssembly(u.get(), makeDevice(opt), t0, A);
        *out = u.release();
    });
}

int symlib_unit_save(symlib_unit_t u, const char* path) {
    if (!u || !path) return 1;
    return guarded([&] { u->unit->save(path); });
}

int symlib_unit_load(symlib_unit_t* out, const char* path, const symlib_options* opt) {
    if (!out || !path) return 1;
    *out = nullptr;
    return guarded([&] {
        std::unique_ptr<symlib_unit> u(new symlib_unit());
        const auto t0 = Clock::now();
        u->unit.reset(new Sym::ComputeUnit<double>(makeDevice(opt), Sym::SavedProgram{path}));
        fillInfo(u.get());
        u->info.build_seconds = seconds(t0, Clock::now());
            *out = u.release();
    });
}

int symlib_unit_compile(symlib_unit_t u) {
    if (!u) return 1;
    return guarded([&] {
        const auto t0 = Clock::now();
        u->unit->compile();
        u->info.compile_seconds = seconds(t0, Clock::now());
    });
}

int symlib_unit_execute(symlib_unit_t u, int num_inputs, const double* const* inputs, const size_t* lengths) {
    if (!u) return 1;
    return guarded([&] {
        if (num_inputs != (int)u->unit->numInputs()) throw std::invalid_argument("wrong number of input containers");
        for (int i = 0; i < num_inputs; ++i) u->unit->setInput(i, inputs[i], lengths[i]);
        u->unit->run(nullptr);
    });
}

int symlib_unit_run(symlib_unit_t u, void* stream) {
    if (!u) return 1;
    return guarded([&] { u->unit->run(stream); });
}

namespace {

Sym::FusedOptions fusedOptionsFrom(const symlib_unit* u, const symlib_options* opt) {
    if (!u->unit->isFused()) throw std::invalid_argument("not a fused-form unit");
    Sym::FusedOptions fo = u->unit->getFusedOptions();
    if (!opt) return fo;
    if (opt->tile_lds_bytes > 0 && (size_t)opt->tile_lds_bytes != fo.ldsBytes)
        throw std::invalid_argument("the tile LDS budget cannot change without rebuilding the unit");
    if (opt->threads_per_block > 0) fo.threadsPerBlock = opt->threads_per_block;
    if (opt->output_groups > 0) fo.outputGroups = opt->output_groups;
    if (opt->waves_per_eu < 0 || opt->waves_per_eu > 8) throw std::invalid_argument("waves per EU must be 0..8");
    fo.wavesPerEU = opt->waves_per_eu;
    if (opt->schedule < 0 || opt->schedule > (int)Sym::FusedSchedule::Strided) throw std::invalid_argument("unknown schedule");
    fo.schedule = (Sym::FusedSchedule)opt->schedule;
    fo.profile = opt->profile_tiles != 0;
    fo.split = opt->split != 0 && u->unit->getFusedProgram().kind == 0;
    fo.ldsDma = opt->register_staging == 0;
    return fo;
}

}  // namespace

int symlib_unit_retune(symlib_unit_t u, const symlib_options* opt) {
    if (!u) return 1;
    return guarded([&] {
        u->unit->retuneFused(fusedOptionsFrom(u, opt));
        u->info.source_bytes = u->unit->getCode().size();
    });
}

const char* symlib_unit_fused_source(symlib_unit_t u, const symlib_options* opt) {
    if (!u) return nullptr;
    const char* out = nullptr;
    int rc = guarded([&] {
        u->scratch = u->unit->fusedSource(fusedOptionsFrom(u, opt));
        out = u->scratch.c_str();
    });
    return rc ? nullptr : out;
}

int symlib_unit_fused_stats(symlib_unit_t u, symlib_fused_stats* st) {
    if (!u || !st) return 1;
    return guarded([&] {
        if (!u->unit->isFused()) throw std::invalid_argument("not a fused-form unit");
        const auto& F = u->unit->getFusedProgram();
        const auto& C = u->unit->getFusedStats();
        const auto& O = u->unit->getFusedOptions();
        *st = symlib_fused_stats{};
        st->num_tiles = F.numTiles;
        st->num_outputs = F.numResults;
        st->num_elements = F.numElements;
        st->element_evaluations = F.elementEvaluations;
        st->contributions = F.numContributions;
        st->lds_bytes = F.ldsBytes;
        st->element_lds_bytes = F.elementLdsBytes;
        st->adds = C.adds;
        st->muls = C.muls;
        st->divs = C.divs;
        st->sqrts = C.sqrts;
        st->statements = C.statements;
        st->max_live = C.maxLive;
        st->output_groups = O.outputGroups;
        st->waves_per_eu = O.wavesPerEU;
        st->threads_per_block = O.threadsPerBlock;
        st->schedule = (int32_t)O.schedule;
        st->profile_tiles = O.profile ? 1 : 0;
    });
}

int symlib_unit_tile_stats(symlib_unit_t u, float* out, size_t n) {
    if (!u) return 1;
    return guarded([&] {
        if (!u->unit->isFused()) throw std::invalid_argument("not a fused-form unit");
        const auto st = Sym::tileStats(u->unit->getFusedProgram());
        if (n != st.size()) throw std::invalid_argument("tile stats: n must be 8 x the number of tiles");
        std::copy(st.begin(), st.end(), out);
    });
}

int symlib_unit_tile_profile(symlib_unit_t u, uint64_t* out, size_t n) {
    if (!u) return 1;
    return guarded([&] {
        const auto st = u->unit->tileProfile();
        if (n != st.size()) throw std::invalid_argument("tile profile: n must be 4 x the number of tiles");
        std::copy(st.begin(), st.end(), out);
    });
}

int symlib_unit_run_graph(symlib_unit_t u, void* stream) {
    if (!u) return 1;
    return guarded([&] { u->unit->runGraph(stream); });
}

int symlib_unit_get_results(symlib_unit_t u, int k, double* out, size_t n) {
    if (!u) return 1;
    return guarded([&] { u->unit->getOutput(k, out, n); });
}

int symlib_unit_num_outputs(symlib_unit_t u) { return u ? (int)u->patterns().size() : 0; }

int symlib_unit_num_inputs(symlib_unit_t u) { return u ? (int)u->unit->numInputs() : 0; }

int symlib_unit_input_length(symlib_unit_t u, int i, uint64_t* n) {
    if (!u || !n || i < 0 || (size_t)i >= u->unit->numInputs()) {
        g_err = "bad input index";
        return 1;
    }
    *n = u->unit->inputSize(i);
    return 0;
}

int symlib_unit_output_shape(symlib_unit_t u, int k, int* rows, int* cols, uint64_t* nnz) {
    if (!u || k < 0 || k >= (int)u->patterns().size()) {
        g_err = "bad output index";
        return 1;
    }
    const auto& P = u->patterns()[k];
    if (rows) *rows = P.rows();
    if (cols) *cols = P.cols();
    if (nnz) *nnz = P.nonZeros();
    return 0;
}

int symlib_unit_output_pattern(symlib_unit_t u, int k, int* outer, int* inner) {
    if (!u || k < 0 || k >= (int)u->patterns().size()) {
        g_err = "bad output index";
        return 1;
    }
    const auto& P = u->patterns()[k];
    if (outer) std::copy(P.outer().begin(), P.outer().end(), outer);
    if (inner) std::copy(P.inner().begin(), P.inner().end(), inner);
    return 0;
}

int symlib_unit_get_info(symlib_unit_t u, symlib_unit_info* info) {
    if (!u || !info) return 1;
    *info = u->info;
    return 0;
}

const char* symlib_unit_source(symlib_unit_t u) { return u ? u->unit->getCode().c_str() : ""; }

symhip_t symlib_unit_runtime(symlib_unit_t u) { return u ? u->unit->runtime() : nullptr; }

void symlib_unit_destroy(symlib_unit_t u) { delete u; }

int symlib_tutorial_known_answers(double out[3]) {
    return guarded([&] {
        using namespace Sym;
        Symbolic a(0, 0), b(1, 0);
        Symbolic c = a + b * sqrt(a * b);
        std::vector<double> vals{1., 2.};
        out[0] = evaluate(c, vals);
        auto dc = differentiate(c, std::vector<Symbolic>{a, b});
        out[1] = evaluate(dc[0], vals);
        out[2] = evaluate(dc[1], vals);
    });
}

}  // extern "C"
