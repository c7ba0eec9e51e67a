// Backward elimination of unknowns from a least-squares solution in one workgroup
// (reference: replacement.py:104-139, a host loop of small tensor operations per step).
//
// W [n, n] float64 (the inverse of the regularized real Gram), b = W q. Deleting
// unknown j raises the objective by b_j^2 / W_jj. Each step removes the cheapest unknown i
// (ties to the lower index) among those still allowed, then keeps the exact inverse of the
// rest through rank-one corrections stored as columns u_s = w_s / sqrt(pivot_s):
//   w = W[:, i] - sum_{s < step} u_s u_s[i],  b -= w b_i / w_i,  d -= u^2,
// where d is the running diagonal. At most `limit` unknowns flagged in `limited` leave.
// 1024 threads; thread t owns rows t + 1024 r (r < kElimRows, so n <= 4096). U [count][n] lives in the
// workspace (coalesced per step); the state b, d, flags stays in registers.
#include "dsplat_internal.h"

namespace dsplat {

constexpr int kElimThreads = 1024;
constexpr int kElimRows = 4;  // n <= 4096
constexpr int kElimMaxCount = 4096;
constexpr int kElimUnroll = 4;

struct Best {
    double cost;
    int index;
};

// NaN first (as an arg-min that propagates NaN), then the smaller cost, then the lower index.
__device__ __forceinline__ bool better(const Best& a, const Best& b) {
    const bool an = isnan(a.cost), bn = isnan(b.cost);
    if (an != bn) return an;
    if (!an && a.cost != b.cost) return a.cost < b.cost;
    return a.index < b.index;
}

__global__ __launch_bounds__(kElimThreads) void elimination_kernel(const double* __restrict__ W,
                                                                   const double* __restrict__ b0, int n, int count,
                                                                   const int* __restrict__ limited, int limit,
                                                                   double* __restrict__ U,
                                                                   long long* __restrict__ removed) {
    __shared__ double u_at[kElimMaxCount];  // u_s[i] of the chosen index, s < step
    __shared__ Best wave_best[kElimThreads / 64];
    __shared__ int chosen;
    __shared__ double chosen_w, chosen_b;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double tiny = 2.2250738585072014e-308;
    double b[kElimRows], d[kElimRows];
    bool alive[kElimRows], lim[kElimRows];
#pragma unroll
    for (int r = 0; r < kElimRows; ++r) {
        const int j = tid + kElimThreads * r;
        const bool in = j < n;
        b[r] = in ? b0[j] : 0.0;
        d[r] = in ? W[(size_t)j * n + j] : 1.0;
        alive[r] = in;
        lim[r] = in && limited ? limited[j] != 0 : false;
    }
    int spent = 0;
    for (int step = 0; step < count; ++step) {
        // arg-min of b^2 / max(d, tiny) over rows that may still leave (blocked rows cost +inf)
        Best best = {INFINITY, n};
#pragma unroll
        for (int r = 0; r < kElimRows; ++r) {
            const int j = tid + kElimThreads * r;
            if (j >= n) break;  // rows of a thread ascend
            const bool blocked = !alive[r] || (limited && lim[r] && spent >= limit);
            const Best c = {blocked ? INFINITY : b[r] * b[r] / fmax(d[r], tiny), j};
            if (better(c, best)) best = c;
        }
        for (int o = 32; o >= 1; o >>= 1) {
            Best other = {__shfl_xor(best.cost, o, 64), __shfl_xor(best.index, o, 64)};
            if (better(other, best)) best = other;
        }
        if (lane == 0) wave_best[wave] = best;
        __syncthreads();
        if (tid == 0) {
            Best all = wave_best[0];
            for (int w = 1; w < kElimThreads / 64; ++w)
                if (better(wave_best[w], all)) all = wave_best[w];
            chosen = all.index < n ? all.index : 0;  // every row blocked: the first, as an arg-min of +inf
        }
        __syncthreads();
        const int i = chosen;
        if (tid < step) u_at[tid] = U[(size_t)tid * n + i];
        for (int s = tid + kElimThreads; s < step; s += kElimThreads) u_at[s] = U[(size_t)s * n + i];
        __syncthreads();
        // w_j = W[j][i] - sum_s U[s][j] u_at[s]; s outermost so each row's chain interleaves
        // with the others' and every U row is one coalesced load per wave
        double w[kElimRows];
#pragma unroll
        for (int r = 0; r < kElimRows; ++r) {
            const int j = tid + kElimThreads * r;
            w[r] = j < n ? W[(size_t)j * n + i] : 0.0;
        }
        int s = 0;
        for (; s + kElimUnroll <= step; s += kElimUnroll) {  // kElimUnroll rows of U in flight
#pragma unroll
            for (int r = 0; r < kElimRows; ++r) {
                const int j = tid + kElimThreads * r;
                if (j < n) {
                    double x[kElimUnroll];
#pragma unroll
                    for (int q = 0; q < kElimUnroll; ++q) x[q] = U[(size_t)(s + q) * n + j];
#pragma unroll
                    for (int q = 0; q < kElimUnroll; ++q) w[r] -= x[q] * u_at[s + q];
                }
            }
        }
        for (; s < step; ++s) {
            const double us = u_at[s];
            const double* __restrict__ row = U + (size_t)s * n;
#pragma unroll
            for (int r = 0; r < kElimRows; ++r) {
                const int j = tid + kElimThreads * r;
                if (j < n) w[r] -= row[j] * us;
            }
        }
#pragma unroll
        for (int r = 0; r < kElimRows; ++r)
            if (tid + kElimThreads * r == i) {
                chosen_w = w[r];
                chosen_b = b[r];
            }
        __syncthreads();
        const double pivot = fmax(chosen_w, tiny);
        const double ratio = chosen_b / pivot, root = sqrt(pivot);
#pragma unroll
        for (int r = 0; r < kElimRows; ++r) {
            const int j = tid + kElimThreads * r;
            if (j >= n) break;
            b[r] -= w[r] * ratio;
            const double u = w[r] / root;
            U[(size_t)step * n + j] = u;
            d[r] -= u * u;
            if (j == i) {
                alive[r] = false;
                d[r] = 1.0;
                b[r] = 0.0;
            }
        }
        if (tid == 0) removed[step] = i;
        if (limited && limited[i] != 0) ++spent;
        __syncthreads();
    }
}

size_t elimination_workspace(int n, int count) {
    if (n < 0 || count < 0 || n > kElimThreads * kElimRows || count > kElimMaxCount || count > n) return 0;
    return sizeof(double) * ((size_t)n * count + 1);
}

hipError_t launch_elimination(const double* W, const double* b, int n, int count, const int* limited, int limit,
                              long long* removed, void* workspace, hipStream_t st) {
    hipLaunchKernelGGL(elimination_kernel, dim3(1), dim3(kElimThreads), 0, st, W, b, n, count, limited, limit,
                       static_cast<double*>(workspace), removed);
    return hipGetLastError();
}

}  // namespace dsplat
