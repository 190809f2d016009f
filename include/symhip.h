/*
 * symhip.h - C ABI of the MI355X runtime that executes compiled expression programs.
 *
 * It replaces the five functions that the reference emits into every generated shared
 * library and resolves with dlsym (src/StringResources.h:198-204, loaded at
 * src/ComputeUnit.cpp:146-152):
 *
 *   reference                                  this ABI
 *   init(indexData, outIndexData, constData)   symhip_create  (tables copied to HBM once)
 *   (system("hipcc ...") + dlopen)             symhip_compile (hipRTC, in process)
 *   setArg(d, i)                               symhip_set_arg / symhip_set_arg_device
 *   run()                                      symhip_run     (all launches on one stream)
 *   getResult(d, k)                            symhip_get_result / symhip_get_result_device
 *   finish()                                   symhip_destroy (frees every device buffer)
 *
 * Differences in behaviour, all deliberate:
 *  - state lives behind an opaque handle instead of file-scope globals, so several programs
 *    can be live in one process;
 *  - every call returns a status code (0 = success, > 0 = hipError_t, < 0 = SYMHIP_E_*)
 *    and leaves a message for symhip_last_error(), instead of exit(code) on a HIP error
 *    (src/StringResources.h:174-181) or a silent no-op when linking failed
 *    (src/ComputeUnit.cpp:155-165);
 *  - results are written in the caller's output order by the program itself: there is no
 *    separate gather kernel (src/StringResources.h:218-221), unless a program in the
 *    reference's own kernel form asks for one (symhip_set_result_gather).
 *
 * Threading: one handle must not be used from two threads at once.  Calls on different
 * handles are independent.
 */
#ifndef SYMHIP_H
#define SYMHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SYMHIP_SUCCESS = 0,
    SYMHIP_E_INVALID = -1,    /* bad argument (null handle, index or size out of range) */
    SYMHIP_E_NO_DEVICE = -2,  /* no HIP device visible */
    SYMHIP_E_COMPILE = -3,    /* hipRTC rejected the source; see symhip_last_error() */
    SYMHIP_E_NOT_COMPILED = -4, /* run before a successful compile / load */
    SYMHIP_E_SYMBOL = -5,     /* a launch names a function the module does not contain */
    SYMHIP_E_IO = -6          /* code-object cache could not be read or written */
};

typedef struct symhip_program* symhip_t;

/* Grid of a persistent launch: as many workgroups as the device keeps resident at once
 * (occupancy of the kernel at its block and LDS size times the compute units). */
#define SYMHIP_GRID_FILL 0xffffffffu

/* One kernel launch of a program; launches run in order on one stream. */
typedef struct {
    const char* name; /* extern "C" __global__ function in the compiled source */
    uint32_t grid;    /* workgroups, or SYMHIP_GRID_FILL */
    uint32_t block;   /* work-items per workgroup */
    uint32_t shared;  /* dynamic LDS bytes per workgroup */
} symhip_launch;

/*
 * Everything the runtime needs besides the source (the reference's generated size arrays,
 * src/ComputeUnit.cpp:376-407, plus its three table files).  All arrays are copied.
 *   values    : FP64 workspace x = [inputs | intermediates], num_values entries
 *   arg_*     : where input container i lives inside x
 *   out_*     : where output container k lives inside the result buffer
 *   index     : u32 gather table (positions inside x)
 *   consts    : FP64 per-instance constants
 *   outpos    : u32 placement table (positions inside the result buffer)
 *   tiles     : u32 tile pool of the tiled and fused forms (may be empty)
 *   work      : u32 words of device scratch, zero at creation (work-queue counters of
 *               persistent launches, which leave them zero when they finish)
 * Generated kernels take (x, x, consts, index, outpos, result, tiles, work) in that order.
 */
typedef struct {
    uint64_t num_values;
    uint32_t num_args;
    const uint64_t* arg_start;
    const uint64_t* arg_length;
    uint32_t num_outputs;
    const uint64_t* out_start;
    const uint64_t* out_length;
    uint64_t num_index;
    const uint32_t* index_data;
    uint64_t num_const;
    const double* const_data;
    uint64_t num_outpos;
    const uint32_t* outpos_data;
    uint64_t num_tiles_data;
    const uint32_t* tiles_data;
    uint32_t num_launches;
    const symhip_launch* launches;
    uint64_t num_work_words;
} symhip_program_desc;

/* 1 when a HIP device is visible, 0 otherwise (never initialises more than the runtime). */
int symhip_device_available(void);

/* Allocate device buffers and upload the tables (reference: init). */
int symhip_create(symhip_t* out, const symhip_program_desc* desc);

/*
 * Compile `source` with hipRTC for the current device's architecture (gfx950 when no device
 * is present) and load it.  `options` is a space-separated list of extra compiler options
 * or NULL.  Code objects are cached on disk by a hash of the source, the full option list
 * (fixed and extra options, architecture) and the hipRTC version.
 */
int symhip_compile(symhip_t h, const char* source, const char* options);

/* Compile without a device and without a handle; writes the code object to `out_path`
 * (used to fill the cache ahead of time).  `arch` e.g. "gfx950". */
int symhip_compile_to_file(const char* source, const char* options, const char* arch, const char* out_path);

/* Compile `source` into the on-disk code-object cache for `arch` (NULL = the current device's,
 * gfx950 without one) unless it is there already, so that a later symhip_compile of the same
 * source and options only loads it.  Thread-safe: several sources may be compiled at once. */
int symhip_cache_compile(const char* source, const char* options, const char* arch);

/* Upload input container i from host memory (reference: setArg, src/StringResources.h:273-275).
 * n must equal its length.  Blocking, like the reference's hipMemcpy: it first waits for all
 * work on the device, so a launch of this program still running on any stream (a non-blocking
 * one included) never sees its inputs change. */
int symhip_set_arg(symhip_t h, int i, const double* host, size_t n);
/* Same, ordered on `stream` (hipStream_t, NULL = default) instead of waiting for the device:
 * the copy runs after the work already queued on that stream and before what follows it.
 * The caller keeps `host` unchanged until the stream has passed the copy. */
int symhip_set_arg_async(symhip_t h, int i, const double* host, size_t n, void* stream);
/* Same from device memory, asynchronously on `stream` (hipStream_t, NULL = default). */
int symhip_set_arg_device(symhip_t h, int i, const double* dev, size_t n, void* stream);

/* Enqueue every launch on `stream` (reference: run()).  Does not synchronise. */
int symhip_run(symhip_t h, void* stream);
/* Same launches replayed from a HIP graph, captured on the first call (the program's
 * launches are the graph's only nodes: no memset or copy nodes).  Later calls only launch
 * the graph on `stream`. */
int symhip_run_graph(symhip_t h, void* stream);
/* Work words of the static scratch array that symhip_run_noalloc uses (64 KiB per device). */
#define SYMHIP_SCRATCH_WORDS 16384u

/*
 * symhip_run without any device allocation.  It has no counterpart in the reference: it is for
 * callers that cannot call hipMalloc (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  The program's launches take their work words (the
 * work-queue counters of persistent schedules, profiling stamps) from a static array of
 * SYMHIP_SCRATCH_WORDS words per device, held until the process ends, instead of the handle's
 * buffer; the call clears the words it uses first (a fill kernel on `stream`, no memset).
 * Programs that need more work words are rejected with SYMHIP_E_INVALID.  Every call on one
 * device shares that device's array: calls that overlap on one device (two streams, two
 * threads) race.  The handle must be compiled; inputs and outputs are the handle's as for
 * symhip_run.  Calls no hipMalloc, hipFree or allocating library routine.
 */
int symhip_run_noalloc(symhip_t h, void* stream);

/* Enqueue launch `idx` only (for per-kernel timing). */
int symhip_launch_one(symhip_t h, uint32_t idx, void* stream);
uint32_t symhip_num_launches(symhip_t h);
/* Workgroups of launch `idx` as launched (SYMHIP_GRID_FILL resolved after compile). */
uint32_t symhip_launch_grid(symhip_t h, uint32_t idx);

/* Re-tuning a compiled program without rebuilding it: overwrite `n` words of the tile pool
 * from word `offset` (after waiting for the device, so no running launch reads them), replace launch `idx` (takes effect at the next symhip_compile), resize
 * the work words (zeroed), read work words back (after synchronising). */
int symhip_update_tiles(symhip_t h, uint64_t offset, uint64_t n, const uint32_t* data);
int symhip_set_launch(symhip_t h, uint32_t idx, const symhip_launch* launch);
/* Replace the whole launch list (its length may change); takes effect at the next compile. */
int symhip_set_launches(symhip_t h, uint32_t n, const symhip_launch* launches);
int symhip_set_work_words(symhip_t h, uint64_t n);
int symhip_read_work(symhip_t h, uint64_t offset, uint64_t n, uint32_t* out);
/*
 * Programs in the reference's kernel form.  The reference emits every group kernel as
 * k<i>(const RealT* x, const RealT* c, RealT* y, const unsigned* p) and launches it as
 * k<i>(x, c + constantsOffset[i], x + outputOffset[i], p + indexOffset[i]) (the K(num) macro,
 * src/StringResources.h:196); its results stay in x and are gathered afterwards as
 * result[j] = x[outIds[j]] (gatherResult, src/StringResources.h:218-221).
 * symhip_set_launch_offsets switches launch `idx` to that four-argument form with these
 * offsets (elements of the constant table, of x and of the index table; each at most its
 * table's length).  symhip_set_launches returns every launch to this runtime's form;
 * symhip_set_launch keeps the form of the launch it replaces.
 * symhip_set_result_gather(h, 1) makes symhip_run, symhip_run_graph and symhip_run_noalloc
 * end with that gather, outpos then holding one position inside x per result (checked here:
 * as many entries as results, each below num_values); 0 turns it off.
 */
int symhip_set_launch_offsets(symhip_t h, uint32_t idx, uint64_t c_offset, uint64_t y_offset, uint64_t p_offset);
int symhip_set_result_gather(symhip_t h, int on);
/* Text to put before the reference's kernels in place of its hipHeader
 * (src/StringResources.h:163-215) when they are compiled with symhip_compile: KERNEL as an
 * extern "C" kernel (hipRTC looks kernels up by unmangled name), RealT = double, INDEX,
 * restrict, and the hipThreadIdx_x / hipBlockDim_x / hipBlockIdx_x names hipRTC lacks. */
const char* symhip_reference_prelude(void);
/* Registers and residency of launch `idx` after compile: VGPRs per work-item, workgroups per
 * compute unit at its block and LDS size, compute units of the device. */
int symhip_launch_info(symhip_t h, uint32_t idx, int* vgprs, int* blocks_per_cu, int* cus);
/* STREAM-style copy of n doubles between 16-byte aligned device buffers, enqueued on `stream`
 * (the measured HBM ceiling; not part of the reference): blocks = 0 copies in one pass, one
 * 16-byte vector per work-item (the fastest form on MI355X); otherwise `blocks` workgroups of
 * 256 work-items loop over the buffers. */
int symhip_stream_copy(double* dst, const double* src, size_t n, uint32_t blocks, void* stream);
/*
 * The implicit-Euler system of a cloth step, the step after the Hessian (reference:
 * examples/BaraffWitkin/implementation/cloth.cpp:97-112, Cloth::generateMatrixAndRHS):
 *   A = M - h (D + h K),   b = M v + h (f - D2 v)
 * for a stiffness K and damping D on one symmetric n x n CSC pattern (outer n + 1, inner nnz,
 * rows sorted per column; every column holds its diagonal entry), the lumped mass m (M =
 * diag(m)), velocities v and forces f (n each).  D = NULL means no damping.  D2 is the
 * symmetric matrix of D's lower triangle (selfadjointView<Lower>), which equals D for the
 * symmetric D taken here.  A is written on K's pattern, b as n values.  Every pointer is device
 * memory (K may be a program's output, symhip_output_device_ptr); enqueued on `stream`.
 */
int symhip_implicit_euler_system(int n, const int* outer, const int* inner, const double* K, const double* D,
                                 const double* m, const double* v, const double* f, double h, double* A, double* b,
                                 void* stream);
/* y = A x for a symmetric n x n CSC A (the product of a conjugate-gradient solve of the system
 * above; column j is read as row j).  Device memory, enqueued on `stream`. */
int symhip_csc_symv(int n, const int* outer, const int* inner, const double* A, const double* x, double* y,
                    void* stream);
/* Rate of the device wall clock read by profiled kernels, in kHz. */
int symhip_wall_clock_khz(int* khz);

/* Copy output container k to host memory after synchronising (reference: getResult). */
int symhip_get_result(symhip_t h, int k, double* host, size_t n);
/* Copy output container k to device memory, asynchronously on `stream`. */
int symhip_get_result_device(symhip_t h, int k, double* dev, size_t n, void* stream);

/* Device address of output container k / of the x workspace (valid until destroy). */
const double* symhip_output_device_ptr(symhip_t h, int k);
double* symhip_values_device_ptr(symhip_t h);

/* Free everything (reference: finish).  Accepts NULL. */
void symhip_destroy(symhip_t h);

/* Message of the last failing call on this thread ("" if none). */
const char* symhip_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SYMHIP_H */
