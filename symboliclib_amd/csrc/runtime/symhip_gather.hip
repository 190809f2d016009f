// Result gather of programs in the reference's form (symhip_set_result_gather): their kernels
// leave every output value in the value buffer x, and the results are collected afterwards as
// result[j] = x[outpos[j]] (the reference's gatherResult, src/StringResources.h:218-221, here
// one launch over all outputs instead of one per output container).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "symhip_internal.hpp"

namespace {

__global__ void __launch_bounds__(256) symhipGatherResults(double* __restrict__ result, const double* __restrict__ x,
                                                           const uint32_t* __restrict__ pos, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * 256u;
    for (uint64_t j = (uint64_t)blockIdx.x * 256u + threadIdx.x; j < n; j += stride) result[j] = x[pos[j]];
}

}  // namespace

int symhipGather(symhip_t h, void* stream) {
    const uint64_t n = h->numResults;
    if (!n) return SYMHIP_SUCCESS;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255u) / 256u, 8192u);
    hipLaunchKernelGGL(symhipGatherResults, dim3(blocks), dim3(256), 0, (hipStream_t)stream, h->result, h->x, h->o, n);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return symhipFail((int)e, std::string("result gather: ") + hipGetErrorString(e));
    return SYMHIP_SUCCESS;
}
