// Bounded reductions and float64 range responses (reference: dsplat/kernels/reduce.slang:1-21,
// response.slang:1-26; bindings dsplat/kernels/reduce.py, response.py).
#include "dsplat_internal.h"

namespace dsplat {

// Per-block sums of |x - y|^2 (wave64 DPP-free shuffle tree + LDS across the 4 waves),
// then one workgroup adds the block sums in float64 in a fixed order. The reference
// reduces 32-lane warps and leaves the block sum to torch (reduce.slang:4-21).
__global__ __launch_bounds__(256) void energy_blocks(const float2* __restrict__ x, const float2* __restrict__ y,
                                                     size_t n, double* __restrict__ part) {
    __shared__ double wsum[4];
    double acc = 0.0;
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const float2 d = make_float2(x[i].x - y[i].x, x[i].y - y[i].y);
        acc += (double)fmaf(d.x, d.x, d.y * d.y);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (wsum[0] + wsum[1]) + (wsum[2] + wsum[3]);
}

__global__ __launch_bounds__(256) void energy_final(const double* __restrict__ part, int blocks,
                                                    float* __restrict__ out) {
    __shared__ double wsum[4];
    double acc = 0.0;
    for (int i = threadIdx.x; i < blocks; i += 256) acc += part[i];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = (float)((wsum[0] + wsum[1]) + (wsum[2] + wsum[3]));
}

int energy_blocks_for(size_t n) {
    size_t b = (n + 255) / 256;
    return (int)(b > kEnergyBlocks ? kEnergyBlocks : (b == 0 ? 1 : b));
}

hipError_t launch_energy(const float2* x, const float2* y, size_t n, double* part, float* out, hipStream_t st) {
    const int blocks = energy_blocks_for(n);
    hipLaunchKernelGGL(energy_blocks, dim3(blocks), dim3(256), 0, st, x, y, n, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(energy_final, dim3(1), dim3(256), 0, st, part, blocks, out);
    return hipGetLastError();
}

// out[bin, peak] = D(bin - center_peak) in float64, centers spread over [low, high]
// (response.slang:2-26): ratio sin(a x) / (S sin(b x)) (Taylor near 0) times e^{-i h x}.
__global__ __launch_bounds__(256) void range_response_kernel(double2* __restrict__ out, double low, double high,
                                                             int count, int bins, int samples, int nfft) {
    const size_t index = blockIdx.x * (size_t)256 + threadIdx.x;
    if (index >= (size_t)count * bins) return;
    const int bin = (int)(index / count), peak = (int)(index % count);
    const double center = count > 1 ? low + (high - low) * (double)peak / (double)(count - 1) : low;
    double x = (double)bin - center;
    x -= (double)nfft * rint(x / (double)nfft);
    const double a = kPiD * samples / nfft, b = kPiD / nfft, h = kPiD * (samples - 1) / nfft;
    double ratio;
    if (fabs(x) < 1e-3) {
        const double c2 = (b * b - a * a) / 6.0;
        const double c4 = (3.0 * a * a * a * a + 7.0 * b * b * b * b - 10.0 * a * a * b * b) / 360.0;
        ratio = 1.0 + c2 * x * x + c4 * x * x * x * x;
    } else {
        ratio = sin(a * x) / ((double)samples * sin(b * x));
    }
    out[index] = make_double2(ratio * cos(h * x), -ratio * sin(h * x));
}

hipError_t launch_range_response(double2* out, double low, double high, int count, int bins, int samples, int nfft,
                                 hipStream_t st) {
    const size_t total = (size_t)count * bins;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(range_response_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, out, low, high,
                       count, bins, samples, nfft);
    return hipGetLastError();
}

}  // namespace dsplat
