// Device-side reductions shared by the kernel files.  Their summation order
// is part of the results: the fused loop kernels (cmbl_loops.hip) replicate
// the plain dot product (k_dot_part + k_dot_final) exactly through them.
#pragma once

#include <hip/hip_runtime.h>

namespace cmbl {

// Sum over the 64 lanes of a wavefront (butterfly; every lane gets the sum).
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Block-wide sum for blockDim.x == NT (multiple of 64): wave sums, then
// thread 0 adds them in wave order starting from 0.  Result valid in thread 0.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sh)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < NT / 64; ++i) t += sh[i];
    }
    __syncthreads();
    return t;
}

}  // namespace cmbl
