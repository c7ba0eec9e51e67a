// Jacobi-preconditioned conjugate gradients of the joint Gauss-Newton system with every
// scalar test on the device (reference: dsplat/varpro.py:206-228, called from
// refinement.py:193-227). The vector algebra of one iteration is one launch of one
// workgroup: no value returns to the host, so the operator and this update can be
// replayed from a HIP graph.
//
// state[0] = rz = r . z, state[1] = 1 while the iteration runs (0 after a stop),
// state[2] = goal = tolerance |rhs|. Vectors are float32, sums float64 in a fixed order.
#include "dsplat_internal.h"

namespace dsplat {

constexpr int kPcgThreads = 1024;

// Workgroup sum in a fixed order: wave trees, then the 16 wave sums by one thread each.
__device__ __forceinline__ double pcg_block_sum(double v, double* scratch) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) scratch[threadIdx.x >> 6] = v;
    __syncthreads();
    double total = 0.0;
#pragma unroll
    for (int w = 0; w < kPcgThreads / 64; ++w) total += scratch[w];
    __syncthreads();  // scratch is reused by the next sum
    return total;
}

// x = 0, r = rhs, d = z = r / diagonal, rz = r . z, goal = tolerance |rhs|.
__global__ __launch_bounds__(kPcgThreads) void pcg_start_kernel(const float* __restrict__ rhs,
                                                                const float* __restrict__ diagonal, int n,
                                                                float tolerance, float* __restrict__ x,
                                                                float* __restrict__ r, float* __restrict__ d,
                                                                float* __restrict__ state) {
    __shared__ double scratch[kPcgThreads / 64];
    double rr = 0.0, rz = 0.0;
    for (int i = threadIdx.x; i < n; i += kPcgThreads) {
        const float ri = rhs[i];
        const float zi = ri / diagonal[i];
        x[i] = 0.0f;
        r[i] = ri;
        d[i] = zi;
        rr += (double)ri * ri;
        rz += (double)ri * zi;
    }
    rr = pcg_block_sum(rr, scratch);
    rz = pcg_block_sum(rz, scratch);
    if (threadIdx.x == 0) {
        state[0] = (float)rz;
        state[1] = 1.0f;
        state[2] = tolerance * (float)sqrt(rr);
        state[3] = 0.0f;
    }
}

// One iteration after image = H d: stop before the update on non-positive (or NaN)
// curvature, else x += step d, r -= step image; stop after it once |r| <= goal; else the
// next direction. A stopped iteration changes nothing.
__global__ __launch_bounds__(kPcgThreads) void pcg_step_kernel(const float* __restrict__ image,
                                                               const float* __restrict__ diagonal, int n,
                                                               float* __restrict__ x, float* __restrict__ r,
                                                               float* __restrict__ d, float* __restrict__ state) {
    __shared__ double scratch[kPcgThreads / 64];
    const float rz = state[0], goal = state[2];
    const bool active = state[1] != 0.0f;
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += kPcgThreads) acc += (double)d[i] * image[i];
    const float curvature = (float)pcg_block_sum(acc, scratch);
    const bool moving = active && curvature > 0.0f;
    const float step = moving ? rz / curvature : 0.0f;
    acc = 0.0;
    for (int i = threadIdx.x; i < n; i += kPcgThreads) {
        float ri = r[i];
        if (moving) {
            x[i] += step * d[i];
            ri -= step * image[i];
            r[i] = ri;
        }
        acc += (double)ri * ri;
    }
    const float norm = (float)sqrt(pcg_block_sum(acc, scratch));
    const bool next = moving && !(norm <= goal);
    acc = 0.0;
    for (int i = threadIdx.x; i < n; i += kPcgThreads) {
        const float ri = r[i];
        acc += (double)ri * (ri / diagonal[i]);
    }
    const float rz_next = (float)pcg_block_sum(acc, scratch);
    if (next) {
        const float beta = rz_next / rz;
        for (int i = threadIdx.x; i < n; i += kPcgThreads) d[i] = r[i] / diagonal[i] + beta * d[i];
    }
    if (threadIdx.x == 0) {  // every thread read state before the first sum's barrier
        state[0] = next ? rz_next : rz;
        state[1] = next ? 1.0f : 0.0f;
    }
}

hipError_t launch_pcg_start(const float* rhs, const float* diagonal, int n, float tolerance, float* x, float* r,
                            float* d, float* state, hipStream_t st) {
    pcg_start_kernel<<<1, kPcgThreads, 0, st>>>(rhs, diagonal, n, tolerance, x, r, d, state);
    return hipGetLastError();
}

hipError_t launch_pcg_step(const float* image, const float* diagonal, int n, float* x, float* r, float* d,
                           float* state, hipStream_t st) {
    pcg_step_kernel<<<1, kPcgThreads, 0, st>>>(image, diagonal, n, x, r, d, state);
    return hipGetLastError();
}

}  // namespace dsplat
