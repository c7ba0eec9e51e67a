// Float64 (and float32) Cholesky factorization and the float64 solve of the K x K VarPro /
// Gauss-Newton systems
// (reference: torch.linalg.cholesky_ex and torch.cholesky_solve in dsplat/varpro.py:44-79,
// which run cuSOLVER potrf / potrs; K is about 2,000-2,500 here).
//
// Right-looking blocked factorization with 64 x 64 blocks, two launches per block column:
//   panel:  every workgroup factors the diagonal block in LDS (the same few microseconds of
//           work in each, which saves a launch), then solves its own block row of the panel,
//           X L_kk^T = A_ik, and zeroes the mirrored block of the upper triangle;
//   update: one workgroup per lower trailing block, A_ij -= L_ik L_jk^T, float64 FMA with
//           a 4 x 4 register tile per thread.
// Only the lower triangle of the input is read. info = 1-based index of the first pivot
// that is not positive (0 if none), as LAPACK potrf reports it, in the factor's own precision
// (the float32 form is the first attempt of varpro.py:62-79); the factor is then not
// meaningful, and the launches of the later block columns return at once. The solve is one workgroup: blocked forward and backward substitution.
#include "dsplat_internal.h"

namespace dsplat {

constexpr int kCholBlock = 64;
constexpr int kCholLd = kCholBlock + 1;  // LDS row stride (doubles): conflict-free columns

// Unblocked Cholesky of the w x w lower block D (LDS, stride kCholLd) by wave 0 alone: lane i
// keeps row i in registers; at step j every lane takes the pivot from LDS, lane i > j scales
// its entry and publishes it in the LDS column buffer, and every lane subtracts
// l_ij l_kj from its entries k > j. One wave's LDS accesses complete in order, so the steps
// need no barrier; the other waves wait at one barrier at the end. Entries above the
// diagonal of a row are updated too and never read. Returns the first failing pivot index
// (or -1) on every thread.
template <typename T>
__device__ int factor_block(T* D, int w) {
    __shared__ int fail;
    __shared__ T col[kCholBlock];
    const int tid = threadIdx.x;
    if (tid < 64) {
        const int i = tid;
        T x[kCholBlock];
#pragma unroll
        for (int k = 0; k < kCholBlock; ++k) x[k] = D[i * kCholLd + k];
        int first = -1;
#pragma unroll
        for (int j = 0; j < kCholBlock; ++j) {
            if (j < w) {  // uniform
                col[i] = x[j];  // column j after the updates of steps < j
                const T d = col[j];
                if (!(d > 0.0) && first < 0) first = j;
                const T ljj = sqrt(d);
                x[j] = i == j ? ljj : x[j] / ljj;
                col[i] = x[j];
#pragma unroll
                for (int k = j + 1; k < kCholBlock; ++k) x[k] = fma(-x[j], col[k], x[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < kCholBlock; ++k) D[i * kCholLd + k] = x[k];
        if (tid == 0) fail = first;
    }
    __syncthreads();
    return fail;
}

template <typename T>
__global__ __launch_bounds__(256) void chol_panel_kernel(T* __restrict__ A, int n, int lda, int kb,
                                                        int* __restrict__ info) {
    __shared__ T D[kCholBlock * kCholLd];
    // an earlier block column failed: the factor is not meaningful, info is final
    if (*info <= n) return;
    const int tid = threadIdx.x;
    const int c0 = kb * kCholBlock, w = min(kCholBlock, n - c0);
    const int ib = kb + blockIdx.x;
    const int r0 = ib * kCholBlock, h = min(kCholBlock, n - r0);
    for (int e = tid; e < kCholBlock * kCholBlock; e += 256) {
        const int i = e >> 6, k = e & 63;
        D[i * kCholLd + k] = (i < w && k <= i) ? A[(size_t)(c0 + i) * lda + c0 + k] : 0.0;
    }
    __syncthreads();
    const int fail = factor_block(D, w);
    if (blockIdx.x == 0) {
        if (tid == 0 && fail >= 0) atomicMin(info, c0 + fail + 1);
        for (int e = tid; e < w * kCholBlock; e += 256) {
            const int i = e >> 6, k = e & 63;
            if (k < w) A[(size_t)(c0 + i) * lda + c0 + k] = k <= i ? D[i * kCholLd + k] : 0.0;
        }
        return;
    }
    // X L^T = B for the h x w block below the diagonal. Rows are independent: the 4 lanes
    // of row i = tid / 4 (one wave holds 16 rows) keep its 16 columns k = tid % 4 + 4 q in
    // registers and pass X[i][j] by a lane shuffle, without barriers.
    const int i = tid >> 2, k0 = tid & 3;
    T x[kCholBlock / 4];
#pragma unroll
    for (int q = 0; q < kCholBlock / 4; ++q) {
        const int k = k0 + 4 * q;
        x[q] = (i < h && k < w) ? A[(size_t)(r0 + i) * lda + c0 + k] : 0.0;
    }
    const int base = (threadIdx.x & 63) & ~3;  // lane of this row's first owner
#pragma unroll
    for (int j = 0; j < kCholBlock; ++j) {
        if (j < w) {  // uniform
            T xj = x[j >> 2] / D[j * kCholLd + j];  // valid on the owner lane k0 == j % 4
            xj = __shfl(xj, base + (j & 3), 64);
            if (k0 == (j & 3)) x[j >> 2] = xj;
#pragma unroll
            for (int q = 0; q < kCholBlock / 4; ++q) {
                const int k = k0 + 4 * q;
                if (k > j) x[q] -= xj * D[k * kCholLd + j];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kCholBlock / 4; ++q) {
        const int k = k0 + 4 * q;
        if (i < h && k < w) {
            A[(size_t)(r0 + i) * lda + c0 + k] = x[q];
            A[(size_t)(c0 + k) * lda + r0 + i] = 0.0;  // mirrored block of the upper triangle
        }
    }
}

// A_ij -= L_ik L_jk^T for the trailing lower blocks kb < jb <= ib.
template <typename T>
__global__ __launch_bounds__(256) void chol_update_kernel(T* __restrict__ A, int n, int lda, int kb,
                                                         const int* __restrict__ info) {
    __shared__ T Li[kCholBlock * (kCholBlock + 1)];  // [p][r]
    __shared__ T Lj[kCholBlock * (kCholBlock + 1)];
    if (*info <= n) return;  // a pivot failed: nothing left to compute
    const int tid = threadIdx.x;
    // linear index -> (ti, tj) with tj <= ti over the trailing blocks
    int t = blockIdx.x, ti = 0;
    while (t > ti) {
        t -= ti + 1;
        ++ti;
    }
    const int tj = t;
    const int c0 = kb * kCholBlock, w = min(kCholBlock, n - c0);
    const int ri = (kb + 1 + ti) * kCholBlock, hi = min(kCholBlock, n - ri);
    const int rj = (kb + 1 + tj) * kCholBlock, hj = min(kCholBlock, n - rj);
    constexpr int LD = kCholBlock + 1;
    for (int e = tid; e < kCholBlock * kCholBlock; e += 256) {
        const int r = e / kCholBlock, p = e % kCholBlock;
        Li[p * LD + r] = (r < hi && p < w) ? A[(size_t)(ri + r) * lda + c0 + p] : 0.0;
        Lj[p * LD + r] = (r < hj && p < w) ? A[(size_t)(rj + r) * lda + c0 + p] : 0.0;
    }
    __syncthreads();
    const int r4 = 4 * (tid >> 4), c4 = 4 * (tid & 15);
    T acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    for (int p = 0; p < w; ++p) {
        T x[4], y[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            x[a] = Li[p * LD + r4 + a];
            y[a] = Lj[p * LD + c4 + a];
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = fma(x[a], y[b], acc[a][b]);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int r = r4 + a, col = c4 + b;
            if (r < hi && col < hj) A[(size_t)(ri + r) * lda + rj + col] -= acc[a][b];
        }
}

// Solve L L^T x = b in place for nrhs columns of B (column j at B + j * ldb), one
// workgroup of 1024 threads; x stays in LDS. Per 64-row block: the product with the solved
// part by all threads, then the 64 x 64 triangle by wave 0 alone (lane = row, the pivot
// value broadcast by a shuffle), so each block costs two barriers.
__global__ __launch_bounds__(1024) void chol_solve_kernel(const double* __restrict__ L, int n, int lda,
                                                         double* __restrict__ B, int nrhs, int ldb) {
    extern __shared__ double xs[];  // [n]
    __shared__ double part[1024];
    __shared__ double T[kCholBlock * kCholLd];  // diagonal block of L
    const int tid = threadIdx.x, row = tid >> 4, sub = tid & 15;
    const int nblk = (n + kCholBlock - 1) / kCholBlock;
    for (int col = 0; col < nrhs; ++col) {
        double* x = B + (size_t)col * ldb;
        for (int i = tid; i < n; i += 1024) xs[i] = x[i];
        __syncthreads();
        // forward: L y = b
        for (int bb = 0; bb < nblk; ++bb) {
            const int r0 = bb * kCholBlock, h = min(kCholBlock, n - r0);
            double s = 0.0;
            if (row < h)
                for (int p = sub; p < r0; p += 16) s = fma(L[(size_t)(r0 + row) * lda + p], xs[p], s);
            part[tid] = s;
            for (int e = tid; e < h * kCholBlock; e += 1024) {
                const int a = e >> 6, b = e & 63;
                if (b < h) T[a * kCholLd + b] = L[(size_t)(r0 + a) * lda + r0 + b];
            }
            __syncthreads();
            if (tid < 64) {
                double y = 0.0;
                if (tid < h) {
                    y = xs[r0 + tid];
                    for (int q = 0; q < 16; ++q) y -= part[tid * 16 + q];
                }
                for (int j = 0; j < h; ++j) {
                    const double yj = __shfl(y / T[j * kCholLd + j], j, 64);
                    if (tid == j) y = yj;
                    if (tid > j && tid < h) y -= T[tid * kCholLd + j] * yj;
                }
                if (tid < h) xs[r0 + tid] = y;
            }
            __syncthreads();
        }
        // backward: L^T x = y
        for (int bb = nblk - 1; bb >= 0; --bb) {
            const int r0 = bb * kCholBlock, h = min(kCholBlock, n - r0);
            double s = 0.0;
            if (row < h)
                for (int p = r0 + h + sub; p < n; p += 16) s = fma(L[(size_t)p * lda + r0 + row], xs[p], s);
            part[tid] = s;
            for (int e = tid; e < h * kCholBlock; e += 1024) {
                const int a = e >> 6, b = e & 63;
                if (b < h) T[a * kCholLd + b] = L[(size_t)(r0 + a) * lda + r0 + b];
            }
            __syncthreads();
            if (tid < 64) {
                double y = 0.0;
                if (tid < h) {
                    y = xs[r0 + tid];
                    for (int q = 0; q < 16; ++q) y -= part[tid * 16 + q];
                }
                for (int j = h - 1; j >= 0; --j) {
                    const double yj = __shfl(y / T[j * kCholLd + j], j, 64);
                    if (tid == j) y = yj;
                    if (tid < j) y -= T[j * kCholLd + tid] * yj;
                }
                if (tid < h) xs[r0 + tid] = y;
            }
            __syncthreads();
        }
        for (int i = tid; i < n; i += 1024) x[i] = xs[i];
        __syncthreads();
    }
}

__global__ void chol_info_reset(int* info, int value) { *info = value; }
__global__ void chol_info_fix(int* info, int none) {
    if (*info == none) *info = 0;
}

template <typename T>
static hipError_t launch_cholesky_t(T* A, int n, int lda, int* info, hipStream_t st) {
    hipLaunchKernelGGL(chol_info_reset, dim3(1), dim3(1), 0, st, info, n + 1);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int nblk = (n + kCholBlock - 1) / kCholBlock;
    for (int kb = 0; kb < nblk; ++kb) {
        hipLaunchKernelGGL(chol_panel_kernel<T>, dim3(nblk - kb), dim3(256), 0, st, A, n, lda, kb, info);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        const int trail = nblk - kb - 1;
        if (trail > 0) {
            hipLaunchKernelGGL(chol_update_kernel<T>, dim3(trail * (trail + 1) / 2), dim3(256), 0, st, A, n, lda,
                               kb, info);
            e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    }
    // n + 1 marks "no failure": map it to 0
    hipLaunchKernelGGL(chol_info_fix, dim3(1), dim3(1), 0, st, info, n + 1);
    return hipGetLastError();
}

hipError_t launch_cholesky(double* A, int n, int lda, int* info, hipStream_t st) {
    return launch_cholesky_t<double>(A, n, lda, info, st);
}

hipError_t launch_cholesky_f32(float* A, int n, int lda, int* info, hipStream_t st) {
    return launch_cholesky_t<float>(A, n, lda, info, st);
}

hipError_t launch_cholesky_solve(const double* L, int n, int lda, double* B, int nrhs, int ldb, hipStream_t st) {
    const size_t smem = sizeof(double) * (size_t)n;
    hipError_t e =
        allow_lds((const void*)chol_solve_kernel, smem + sizeof(double) * (1024 + kCholBlock * kCholLd));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(chol_solve_kernel, dim3(1), dim3(1024), smem, st, L, n, lda, B, nrhs, ldb);
    return hipGetLastError();
}

}  // namespace dsplat
