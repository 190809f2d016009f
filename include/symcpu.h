/*
 * symcpu.h - C ABI that runs a compiled expression program on the host CPU.
 *
 * It restates the reference's CPU path of ComputeUnit (Device(VecWidth(w), NumThreads(n)),
 * src/ComputeUnit.cpp:121-133 and :409-447): the program's straight-line kernels are emitted
 * as host C++ (loops over instances split across OpenMP threads), compiled by a host clang
 * with the reference's flags into a shared library and bound with dlopen/dlsym.  The tables
 * are those of the GPU waves form (symhip.h's symhip_program_desc; launches are ignored).
 *
 *   reference (CPU device)                        this ABI
 *   system("clang++ -O3 -ffast-math ...") + dlopen symcpu_compile (library cached on disk)
 *   setArg(d, i)                                  symcpu_set_arg
 *   run()                                         symcpu_run(h, threads)
 *   getResult(d, k)                               symcpu_get_result
 *
 * Compiler: $SYMLIB_CPU_CXX, else /opt/rocm/llvm/bin/clang++.  Flags (reference,
 * src/ComputeUnit.cpp:125, :132): -O3 -ffast-math -fno-trapping-math -fno-math-errno
 * -fno-signed-zeros -std=c++17 -msse4.2 -mavx2 -mfma -fopenmp; extra flags may be appended.
 * Status codes are those of symhip.h.
 */
#ifndef SYMCPU_H
#define SYMCPU_H

#include <stddef.h>

#include "symhip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct symcpu_program* symcpu_t;

/* Host copies of the tables and buffers of `desc` (its launches are ignored). */
int symcpu_create(symcpu_t* out, const symhip_program_desc* desc);
/* Compile `source` (emitCpu's text: extern "C" symlib_cpu_run) into a shared library in the
 * code cache and load it; `extra_flags` (NULL = none) are appended to the reference flags. */
int symcpu_compile(symcpu_t h, const char* source, const char* extra_flags);
/* Copy input container i (n values) into the value buffer. */
int symcpu_set_arg(symcpu_t h, int i, const double* host, size_t n);
/* Evaluate every kernel with `threads` OpenMP threads (>= 1). */
int symcpu_run(symcpu_t h, int threads);
/* Copy output container k (n values). */
int symcpu_get_result(symcpu_t h, int k, double* host, size_t n);
void symcpu_destroy(symcpu_t h);
const char* symcpu_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* SYMCPU_H */
