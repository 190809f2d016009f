// symhip_stream_copy: a STREAM-style device copy, the measured HBM ceiling that kernel
// throughputs are compared with besides the nominal 8 TB/s.  Two forms: one pass (every
// work-item moves one 16-byte vector, as many workgroups as that takes: the most requests in
// flight the dispatcher can keep resident, the fastest form on MI355X), or a grid-stride loop
// over a given number of workgroups in which each work-item moves 64 bytes per iteration as
// four 16-byte loads issued before four 16-byte stores.  Stores are non-temporal (the copy's
// destination is not read again).
#include <hip/hip_runtime.h>

#include <string>

#include "symhip_internal.hpp"

namespace {

typedef double v2d __attribute__((ext_vector_type(2)));

__global__ void __launch_bounds__(256) symhipStreamCopy(v2d* __restrict__ dst, const v2d* __restrict__ src, size_t n2) {
    const size_t stride = (size_t)gridDim.x * 256u;
    size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    for (; i + 3 * stride < n2; i += 4 * stride) {
        const v2d a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
        __builtin_nontemporal_store(a, dst + i);
        __builtin_nontemporal_store(b, dst + i + stride);
        __builtin_nontemporal_store(c, dst + i + 2 * stride);
        __builtin_nontemporal_store(d, dst + i + 3 * stride);
    }
    for (; i < n2; i += stride) __builtin_nontemporal_store(src[i], dst + i);
}

__global__ void __launch_bounds__(256) symhipStreamCopyOnePass(v2d* __restrict__ dst, const v2d* __restrict__ src,
                                                                size_t n2) {
    const size_t i = (size_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n2) __builtin_nontemporal_store(src[i], dst + i);
}

__global__ void __launch_bounds__(64) symhipStreamTail(double* __restrict__ dst, const double* __restrict__ src) {
    if (threadIdx.x == 0) dst[0] = src[0];
}

}  // namespace

extern "C" int symhip_stream_copy(double* dst, const double* src, size_t n, uint32_t blocks, void* stream) {
    if (n && (!dst || !src)) return symhipFail(SYMHIP_E_INVALID, "symhip_stream_copy: bad argument");
    if (((uintptr_t)dst | (uintptr_t)src) & 15u)
        return symhipFail(SYMHIP_E_INVALID, "symhip_stream_copy: buffers must be 16-byte aligned");
    const size_t n2 = n / 2;
    if (blocks == 0 && (n2 + 255u) / 256u > 0x7fffffffu)
        return symhipFail(SYMHIP_E_INVALID, "symhip_stream_copy: too long for one pass");
    if (n2 && blocks == 0)
        hipLaunchKernelGGL(symhipStreamCopyOnePass, dim3((uint32_t)((n2 + 255u) / 256u)), dim3(256), 0,
                           (hipStream_t)stream, (v2d*)dst, (const v2d*)src, n2);
    else if (n2)
        hipLaunchKernelGGL(symhipStreamCopy, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (v2d*)dst,
                           (const v2d*)src, n2);
    if (n & 1) hipLaunchKernelGGL(symhipStreamTail, dim3(1), dim3(64), 0, (hipStream_t)stream, dst + n - 1, src + n - 1);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return symhipFail((int)e, std::string("symhip_stream_copy: ") + hipGetErrorString(e));
    return SYMHIP_SUCCESS;
}
