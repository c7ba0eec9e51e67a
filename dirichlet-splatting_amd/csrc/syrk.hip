// Symmetric rank-k update on FP32 MFMA: upper(out) += rows^T rows (reference: the cuBLAS
// SSYRK call of dsplat/kernels/cublas.py:31-44; the lower triangle is left untouched).
//
// rows [m, n] row-major, out [n, n] row-major. A workgroup (4 waves) owns a 64 x 64
// output tile of the upper triangle; K = m is walked in chunks of 32 rows staged in LDS,
// each wave accumulating a 32 x 32 block with v_mfma_f32_32x32x2f32. Lane layout
// (gfx950): A[i][k] -> lane (i + 32 k), B[k][j] -> lane (j + 32 k),
// C[8 (r / 4) + 4 (lane / 32) + r % 4][lane % 32] -> acc[r].
#include "dsplat_internal.h"

namespace dsplat {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kSyrkTile = 64;
constexpr int kSyrkChunk = 32;

__device__ __forceinline__ void tile_of(int t, int nb, int* bi, int* bj) {
    // t-th tile of the upper triangle, row by row: row bi holds tiles bi..nb-1
    int row = 0;
    while (t >= nb - row) {
        t -= nb - row;
        ++row;
    }
    *bi = row;
    *bj = row + t;
}

__global__ __launch_bounds__(256) void syrk_kernel(const float* __restrict__ rows, int m, int n,
                                                   float* __restrict__ out) {
    __shared__ float Xi[kSyrkChunk][kSyrkTile + 4];
    __shared__ float Xj[kSyrkChunk][kSyrkTile + 4];
    const int nb = (n + kSyrkTile - 1) / kSyrkTile;
    int bi, bj;
    tile_of(blockIdx.x, nb, &bi, &bj);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wi = wave >> 1, wj = wave & 1;
    const int ci = bi * kSyrkTile, cj = bj * kSyrkTile;
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    for (int k0 = 0; k0 < m; k0 += kSyrkChunk) {
        __syncthreads();
        for (int e = tid; e < kSyrkChunk * kSyrkTile; e += 256) {
            const int kr = e / kSyrkTile, c = e - kr * kSyrkTile;
            const int k = k0 + kr;
            const bool krow = k < m;
            Xi[kr][c] = (krow && ci + c < n) ? rows[(size_t)k * n + ci + c] : 0.0f;
            Xj[kr][c] = (krow && cj + c < n) ? rows[(size_t)k * n + cj + c] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kSyrkChunk; kk += 2) {
            const float a = Xi[kk + (lane >> 5)][wi * 32 + (lane & 31)];
            const float b = Xj[kk + (lane >> 5)][wj * 32 + (lane & 31)];
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc, 0, 0, 0);
        }
    }
    const int col = cj + wj * 32 + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = ci + wi * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (row < n && col < n && row <= col) out[(size_t)row * n + col] += acc[r];
    }
}

hipError_t launch_syrk(const float* rows, int m, int n, float* out, hipStream_t st) {
    const int nb = (n + kSyrkTile - 1) / kSyrkTile;
    const int tiles = nb * (nb + 1) / 2;
    if (tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(syrk_kernel, dim3(tiles), dim3(256), 0, st, rows, m, n, out);
    return hipGetLastError();
}

}  // namespace dsplat
