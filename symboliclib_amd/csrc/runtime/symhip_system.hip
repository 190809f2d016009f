// The step after the Hessian: the implicit-Euler system of a cloth time step and the products a
// solver of it needs (examples/BaraffWitkin/implementation/cloth.cpp:97-112 of the reference:
// A = M - h (D + h K), b = M v + h (f - D2 v), with D2 the symmetric matrix of D's lower
// triangle), on the device, reading the stiffness K where the compiled program left it.
//
// Layout: one symmetric n x n CSC pattern (outer n + 1, inner nnz, rows sorted per column) for
// K, the damping D and the system A; the lumped mass m, velocities v, forces f and every vector
// are n doubles.  Symmetry lets column j stand for row j, so every product is a gather over the
// column's entries and every output is written once (no atomics).
//
// Work split: SEG lanes per column (a wave covers 64 / SEG consecutive columns, whose entries
// are contiguous in memory, so the value and index reads of a wave are one contiguous range);
// a lane strides over its column's entries and the column's sum is reduced across its SEG
// lanes with cross-lane shuffles.  HBM-bound: per nonzero the assembly reads K, D and the row
// index and writes A (28 B, 20 B without damping); the symmetric product reads A and the row
// index (12 B); the gathered vector entries come from the caches.
#include <hip/hip_runtime.h>

#include <string>

#include "symhip_internal.hpp"

namespace {

constexpr int SEG = 16;
constexpr int BLOCK = 256;

__device__ __forceinline__ double segmentSum(double s) {
#pragma unroll
    for (int o = SEG / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, SEG);
    return s;
}

// A = M - h (D + h K) on K's pattern (M on the diagonal entries); b = M v + h (f - D v).
template <bool DAMPING>
__global__ void __launch_bounds__(BLOCK) symhipImplicitEulerSystem(int n, const int* __restrict__ outer,
                                                                  const int* __restrict__ inner,
                                                                  const double* __restrict__ K,
                                                                  const double* __restrict__ D,
                                                                  const double* __restrict__ m,
                                                                  const double* __restrict__ v,
                                                                  const double* __restrict__ f, double h,
                                                                  double* __restrict__ A, double* __restrict__ b) {
    const int j = (int)((blockIdx.x * (unsigned)BLOCK + threadIdx.x) / SEG);
    const int lane = (int)(threadIdx.x % SEG);
    if (j >= n) return;  // whole segments leave together: a segment never straddles this test
    const int p0 = outer[j], p1 = outer[j + 1];
    const double mj = m[j];
    double dv = 0.0;
    for (int p = p0 + lane; p < p1; p += SEG) {
        const int i = inner[p];
        const double d = DAMPING ? D[p] : 0.0;
        const double a = -h * (d + h * K[p]);
        A[p] = i == j ? mj + a : a;
        if (DAMPING) dv += d * v[i];
    }
    if (DAMPING) dv = segmentSum(dv);
    if (lane == 0) b[j] = mj * v[j] + h * (f[j] - dv);
}

// y = A x for a symmetric A (column j read as row j).
__global__ void __launch_bounds__(BLOCK) symhipCscSymv(int n, const int* __restrict__ outer,
                                                       const int* __restrict__ inner,
                                                       const double* __restrict__ A, const double* __restrict__ x,
                                                       double* __restrict__ y) {
    const int j = (int)((blockIdx.x * (unsigned)BLOCK + threadIdx.x) / SEG);
    const int lane = (int)(threadIdx.x % SEG);
    if (j >= n) return;
    const int p0 = outer[j], p1 = outer[j + 1];
    double s = 0.0;
    for (int p = p0 + lane; p < p1; p += SEG) s += A[p] * x[inner[p]];
    s = segmentSum(s);
    if (lane == 0) y[j] = s;
}

uint32_t gridFor(int n) { return (uint32_t)(((size_t)n * SEG + BLOCK - 1) / BLOCK); }

int launched(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return symhipFail((int)e, std::string(what) + ": " + hipGetErrorString(e));
    return SYMHIP_SUCCESS;
}

}  // namespace

extern "C" int symhip_implicit_euler_system(int n, const int* outer, const int* inner, const double* K,
                                            const double* D, const double* m, const double* v, const double* f,
                                            double h, double* A, double* b, void* stream) {
    if (n < 0 || (n > 0 && (!outer || !inner || !K || !m || !v || !f || !A || !b)))
        return symhipFail(SYMHIP_E_INVALID, "symhip_implicit_euler_system: bad argument");
    if (n == 0) return SYMHIP_SUCCESS;
    if ((size_t)n * SEG > 0x7fffffffull * BLOCK)
        return symhipFail(SYMHIP_E_INVALID, "symhip_implicit_euler_system: too many columns");
    if (D)
        hipLaunchKernelGGL(symhipImplicitEulerSystem<true>, dim3(gridFor(n)), dim3(BLOCK), 0, (hipStream_t)stream, n,
                           outer, inner, K, D, m, v, f, h, A, b);
    else
        hipLaunchKernelGGL(symhipImplicitEulerSystem<false>, dim3(gridFor(n)), dim3(BLOCK), 0, (hipStream_t)stream, n,
                           outer, inner, K, D, m, v, f, h, A, b);
    return launched("symhip_implicit_euler_system");
}

extern "C" int symhip_csc_symv(int n, const int* outer, const int* inner, const double* A, const double* x, double* y,
                               void* stream) {
    if (n < 0 || (n > 0 && (!outer || !inner || !A || !x || !y)))
        return symhipFail(SYMHIP_E_INVALID, "symhip_csc_symv: bad argument");
    if (n == 0) return SYMHIP_SUCCESS;
    if ((size_t)n * SEG > 0x7fffffffull * BLOCK) return symhipFail(SYMHIP_E_INVALID, "symhip_csc_symv: too many columns");
    hipLaunchKernelGGL(symhipCscSymv, dim3(gridFor(n)), dim3(BLOCK), 0, (hipStream_t)stream, n, outer, inner, A, x, y);
    return launched("symhip_csc_symv");
}
