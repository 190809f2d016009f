/*
 * symlib.h - C ABI over the C++ ComputeUnit for the three benchmark workloads, for callers
 * that cannot instantiate the C++ templates (Python ctypes, other FFIs).
 *
 * Each symlib_unit_* constructor runs the workload's symbolic front-end exactly as the
 * reference's example programs do and then constructs
 *     Sym::ComputeUnit<double>(Sym::Device(Sym::UseHIP(), Sym::ThreadsPerBlock(t)), in, out)
 * (reference surface: src/ComputeUnit.hpp:192-222):
 *   symlib_unit_ata_plus_a   examples/Tutorial/main.cpp:36-68     in: A values, out: A^T A + A
 *   symlib_unit_cotan        examples/BuildCotan/main.cpp:39-121  in: V (n x 3), out: L
 *   symlib_unit_bw_hessian   examples/BaraffWitkinAD/main.cpp:88-113 (+ bending stiffness)
 *                                                                  in: X (n x 3), out: H
 * symlib_unit_compile / _execute / _get_results are ComputeUnit::compile / execute /
 * getResults.  Dense n x 3 inputs are column-major (x of all vertices, then y, then z).
 * Output containers are CSC matrices; their pattern is available before compilation.
 *
 * All functions return 0 on success and a non-zero code otherwise; symlib_last_error()
 * then holds the message.  No function throws across the ABI.
 */
#ifndef SYMLIB_H
#define SYMLIB_H

#include <stddef.h>
#include <stdint.h>

#include "symhip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct symlib_unit* symlib_unit_t;

typedef struct {
    int threads_per_block;       /* 0 = default (256) */
    int decomposition_threshold; /* < 0 = default (6) */
    int simplify;                /* 1 = simplify kernel templates (default), 0 = skip */
    int share_templates;         /* 1 = one differentiated template per element shape (default) */
    int tiled;                   /* 1 = LDS-staged work tiles when possible (default), 0 = wave launches */
    int tile_lds_bytes;          /* LDS bytes per workgroup, 0 = default (tiled form 64 KiB; fused
                                    forms 53760, all regions, so three workgroups share a CU) */
    int direct_assembly;         /* cloth Hessian: 1 = HessianAssembly (no per-nonzero expressions),
                                    0 = symbolic SparseMatrix through ComputeUnit (default) */
    int fused;                   /* cloth Hessian: 1 = fused assembly form (implies direct_assembly):
                                    tiles of owner vertices, element Hessians in LDS, one launch */
    int schedule;                /* fused form: 0 one workgroup per tile, 1 persistent work queue
                                    (default), 2 queue with largest tiles first, 3 queue with
                                    instances interleaved, 4 persistent strided */
    int output_groups;           /* fused form: work items per element, 0 = default (5) */
    int waves_per_eu;            /* fused form: minimum waves per SIMD to compile for, 0 = none */
    int profile_tiles;           /* fused form: record wall-clock stamps per tile (symlib_unit_tile_profile) */
    int split;                   /* fused cloth form: 1 = two launches (element images in HBM, then sums) */
    int register_staging;        /* fused forms: 1 = stage tiles through registers instead of
                                    direct global-to-LDS loads (default 0) */
    int output_sums;             /* fused element tiles: one contribution list per output even
                                    when the Hessian is made of node blocks (block sums otherwise) */
    int cpu_threads;             /* > 0: the reference's CPU device instead of UseHIP, with this
                                    many OpenMP threads (waves form only; symcpu.h); 0 = GPU */
} symlib_options;

/* Fused-form program and code statistics (the autotuner's features). */
typedef struct {
    uint64_t num_tiles, num_outputs, num_elements, element_evaluations, contributions;
    uint64_t lds_bytes, element_lds_bytes;         /* per workgroup: all LDS, element part */
    uint64_t adds, muls, divs, sqrts, statements;  /* per element, over all output groups */
    int32_t max_live;                              /* most values a work item keeps live */
    int32_t output_groups, waves_per_eu, threads_per_block, schedule, profile_tiles;
} symlib_fused_stats;

typedef struct {
    uint64_t num_inputs;        /* input values */
    uint64_t num_intermediates; /* FP64 intermediates written and re-read */
    uint64_t num_index;         /* u32 gather entries */
    uint64_t num_const;         /* FP64 per-instance constants */
    uint64_t num_outpos;        /* u32 placement entries */
    uint64_t num_results;       /* output values */
    uint32_t num_kernels;
    uint32_t num_waves;
    uint64_t source_bytes;
    double frontend_seconds;    /* symbolic front-end (building the expressions) */
    double build_seconds;       /* ComputeUnit construction (decompose .. emit) */
    double compile_seconds;     /* hipRTC + table upload */
    uint32_t tiled;             /* 0 = wave launches, 1 = tile-fused form, 2 = fused assembly form */
    uint64_t tile_pool_entries; /* u32 entries of the tile tables */
    char tiling[128];           /* tile count, LDS bytes and recomputation factor, or why untiled */
} symlib_unit_info;

void symlib_default_options(symlib_options* opt);

int symlib_unit_ata_plus_a(symlib_unit_t* out, int n, const int* colptr, const int* rowidx, const symlib_options* opt);
int symlib_unit_cotan(symlib_unit_t* out, int num_vertices, int num_faces, const int* faces, const symlib_options* opt);
int symlib_unit_bw_hessian(symlib_unit_t* out, int num_vertices, int num_faces, const int* faces, const double* rest,
                           double k_stretch, double k_shear, double k_bend, const symlib_options* opt);

/* Batched cloth Hessians: `count` independent cloths in one program (always directly
 * assembled).  Cloth k has nv[k] vertices, nf[k] faces faces[k] (3 ints per face) and rest
 * positions rest[k] (column-major nv x 3); it is input container k and output container k. */
int symlib_unit_bw_hessian_batch(symlib_unit_t* out, int count, const int* num_vertices, const int* num_faces,
                                 const int* const* faces, const double* const* rest, double k_stretch, double k_shear,
                                 double k_bend, const symlib_options* opt);

/* The built program on disk (reference: the constData / indexData / outIndexData files init()
 * writes and the generated library reads, src/ComputeUnit.cpp:364-366,
 * src/StringResources.h:233-269): save writes the unit's tables, templates, source, launches,
 * tile pools and output patterns to one file; load constructs the same unit from it without
 * front-end or build (compile then finds the code object in the runtime's cache).  opt picks
 * the device kind (cpu_threads > 0: the CPU device), which must be the saved one; every other
 * setting comes from the file.  Files are checked on load (magic, version, sizes, indices). */
int symlib_unit_save(symlib_unit_t u, const char* path);
int symlib_unit_load(symlib_unit_t* out, const char* path, const symlib_options* opt);

int symlib_unit_compile(symlib_unit_t u);
/* inputs[i] has lengths[i] values; one entry per input container, in order */
int symlib_unit_execute(symlib_unit_t u, int num_inputs, const double* const* inputs, const size_t* lengths);
/* enqueue the program on a stream without copying inputs (they are already on the device) */
int symlib_unit_run(symlib_unit_t u, void* stream);
/* same, replayed from a HIP graph captured on the first call */
int symlib_unit_run_graph(symlib_unit_t u, void* stream);
int symlib_unit_get_results(symlib_unit_t u, int k, double* out, size_t n);

int symlib_unit_num_outputs(symlib_unit_t u);
/* input containers and the number of values of input container i */
int symlib_unit_num_inputs(symlib_unit_t u);
int symlib_unit_input_length(symlib_unit_t u, int i, uint64_t* n);
/* rows, cols, nnz of output container k */
int symlib_unit_output_shape(symlib_unit_t u, int k, int* rows, int* cols, uint64_t* nnz);
/* CSC outer (cols + 1) and inner (nnz) arrays of output container k */
int symlib_unit_output_pattern(symlib_unit_t u, int k, int* outer, int* inner);
int symlib_unit_get_info(symlib_unit_t u, symlib_unit_info* info);

/* Fused form only.  retune: apply output_groups, waves_per_eu, threads_per_block, schedule and
 * profile_tiles without rebuilding the tiles (tile_lds_bytes must be 0 or unchanged) and
 * recompile if compiled; fused_source: the source those options would emit (valid until the
 * next call on this unit); tile_stats: TILE_STATS (8) floats per tile in launch order (owned
 * outputs, element contributions, elements, staged doubles, element LDS values, recipe
 * words, output constants, runs); tile_profile: 4 wall-clock stamps per tile in launch order
 * (start, staging done, elements done, outputs written) of the last profiled launch. */
int symlib_unit_retune(symlib_unit_t u, const symlib_options* opt);
const char* symlib_unit_fused_source(symlib_unit_t u, const symlib_options* opt);
int symlib_unit_fused_stats(symlib_unit_t u, symlib_fused_stats* stats);
int symlib_unit_tile_stats(symlib_unit_t u, float* out, size_t n);
int symlib_unit_tile_profile(symlib_unit_t u, uint64_t* out, size_t n);
/* emitted HIP source (valid until destroy) */
const char* symlib_unit_source(symlib_unit_t u);
/* runtime handle for device-side access (NULL before compile) */
symhip_t symlib_unit_runtime(symlib_unit_t u);
void symlib_unit_destroy(symlib_unit_t u);

/* Known answers of examples/Tutorial/main.cpp:22-33: evaluate(a + b sqrt(ab)) and both
 * derivatives at (a, b) = (1, 2); host only. */
int symlib_tutorial_known_answers(double out[3]);

const char* symlib_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SYMLIB_H */
