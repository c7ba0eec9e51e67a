// Backward elimination of unknowns from a least-squares solution (reference:
// replacement.py:104-139, a host loop of small tensor operations per step).
//
// W [n, n] float64 (the inverse of the regularized real Gram), b = W q. Deleting
// unknown j raises the objective by b_j^2 / W_jj. Each step removes the cheapest unknown i
// (ties to the lower index) among those still allowed, then keeps the exact inverse of the
// rest through rank-one corrections stored as rows u_s = w_s / sqrt(pivot_s) of U:
//   w = W[:, i] - sum_{s < step} u_s u_s[i],  b -= w b_i / w_i,  d -= u^2,
// where d is the running diagonal. At most `limit` unknowns flagged in `limited` leave.
//
// One launch per step over workgroups that each own 64 rows (256 threads: 4 threads per row
// split the sum over s), so the O(n step) gather of w is spread over the chip. The state
// (b, d, alive, the number of flagged unknowns gone) is double-buffered in the workspace:
// step k reads buffer k % 2 and writes the other, and every workgroup reduces the per-group
// arg-min candidates of the previous launch itself, so no step needs a grid-wide barrier.
// Every workgroup computes the pivot w_i = W[i][i] - sum_s u_s[i]^2 from the gathered u_s[i]
// in the same fixed order, so all of them apply the same update.
#include "dsplat_internal.h"

namespace dsplat {

constexpr int kElimRowsPerGroup = 64;
constexpr int kElimSlices = 4;  // threads per row
constexpr int kElimThreads = kElimRowsPerGroup * kElimSlices;
constexpr int kElimMaxCount = 4096;

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

__device__ __forceinline__ Best wave_best(Best best) {
    for (int o = 32; o >= 1; o >>= 1) {
        const Best other = {__shfl_xor(best.cost, o, 64), __shfl_xor(best.index, o, 64)};
        if (better(other, best)) best = other;
    }
    return best;
}

struct ElimState {
    double* U;      // [count][n]
    double* b[2];   // [n]
    double* d[2];   // [n]
    int* alive[2];  // [n]
    Best* best[2];  // [groups] arg-min candidates of the next step
    int* spent;     // [2]
};

static ElimState elim_layout(void* ws, int n, int count, int groups) {
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    char* p = static_cast<char*>(ws);
    ElimState s;
    s.U = reinterpret_cast<double*>(p);
    p += up(sizeof(double) * (size_t)n * count);
    for (int k = 0; k < 2; ++k) {
        s.b[k] = reinterpret_cast<double*>(p);
        p += up(sizeof(double) * (size_t)n);
        s.d[k] = reinterpret_cast<double*>(p);
        p += up(sizeof(double) * (size_t)n);
        s.alive[k] = reinterpret_cast<int*>(p);
        p += up(sizeof(int) * (size_t)n);
        s.best[k] = reinterpret_cast<Best*>(p);
        p += up(sizeof(Best) * (size_t)groups);
    }
    s.spent = reinterpret_cast<int*>(p);
    return s;
}

static size_t elim_bytes(int n, int count, int groups) {
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    return up(sizeof(double) * (size_t)n * count) +
           2 * (2 * up(sizeof(double) * (size_t)n) + up(sizeof(int) * (size_t)n) + up(sizeof(Best) * (size_t)groups)) +
           up(2 * sizeof(int));
}

// Arg-min of b^2 / max(d, tiny) over this group's rows that may still leave (blocked rows cost
// +inf), written as the group's candidate. Threads of slice 0 hold the rows.
__device__ void group_candidate(double bj, double dj, bool allowed, int j, bool row_thread, Best* shared_best,
                                Best* out) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const double tiny = 2.2250738585072014e-308;
    Best c = {INFINITY, 0x7fffffff};
    if (row_thread) c = {allowed ? bj * bj / fmax(dj, tiny) : INFINITY, j};
    c = wave_best(c);
    if (lane == 0) shared_best[wave] = c;
    __syncthreads();
    if (tid == 0) {
        Best all = shared_best[0];
        for (int w = 1; w < kElimThreads / 64; ++w)
            if (better(shared_best[w], all)) all = shared_best[w];
        *out = all;
    }
}

__global__ __launch_bounds__(kElimThreads) void elim_init_kernel(const double* __restrict__ W,
                                                                 const double* __restrict__ b0, int n,
                                                                 const int* __restrict__ limited, int limit,
                                                                 ElimState st) {
    __shared__ Best shared_best[kElimThreads / 64];
    const int tid = threadIdx.x;
    const int j = blockIdx.x * kElimRowsPerGroup + (tid & (kElimRowsPerGroup - 1));
    const bool row_thread = tid < kElimRowsPerGroup && j < n;
    double bj = 0.0, dj = 1.0;
    if (row_thread) {
        bj = b0[j];
        dj = W[(size_t)j * n + j];
        st.b[0][j] = bj;
        st.d[0][j] = dj;
        st.alive[0][j] = 1;
    }
    const bool allowed = !(limited && row_thread && limited[j] != 0 && 0 >= limit);
    group_candidate(bj, dj, allowed, j, row_thread, shared_best, st.best[0] + blockIdx.x);
    if (blockIdx.x == 0 && tid == 0) st.spent[0] = 0;
}

__global__ __launch_bounds__(kElimThreads) void elim_step_kernel(const double* __restrict__ W, int n, int step,
                                                                 const int* __restrict__ limited, int limit,
                                                                 ElimState st, long long* __restrict__ removed) {
    __shared__ double u_at[kElimMaxCount];  // u_s[i] of the chosen index, s < step
    __shared__ double part[kElimSlices][kElimRowsPerGroup];
    __shared__ double red[kElimThreads];
    __shared__ Best shared_best[kElimThreads / 64];
    __shared__ int chosen;
    const int tid = threadIdx.x, lane = tid & 63;
    const int in = step & 1, out = in ^ 1, groups = gridDim.x;
    // the step's unknown: the best of the previous launch's group candidates
    if (tid < 64) {
        Best best = {INFINITY, 0x7fffffff};
        for (int g = lane; g < groups; g += 64)
            if (better(st.best[in][g], best)) best = st.best[in][g];
        best = wave_best(best);
        if (tid == 0) chosen = best.index < n ? best.index : 0;  // every row blocked: the first
    }
    __syncthreads();
    const int i = chosen;
    for (int s = tid; s < step; s += kElimThreads) u_at[s] = st.U[(size_t)s * n + i];
    __syncthreads();
    // pivot = W[i][i] - sum_s u_s[i]^2, the same fixed-order sum in every workgroup
    double acc = 0.0;
    for (int s = tid; s < step; s += kElimThreads) acc = fma(u_at[s], u_at[s], acc);
    red[tid] = acc;
    __syncthreads();
    for (int w = kElimThreads / 2; w >= 1; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    const double tiny = 2.2250738585072014e-308;
    const double pivot = fmax(W[(size_t)i * n + i] - red[0], tiny);
    const double bi = st.b[in][i];
    // w_j for this group's rows: slice q of row r sums s = q, q + 4, ...
    const int r = tid & (kElimRowsPerGroup - 1), q = tid / kElimRowsPerGroup;
    const int j = blockIdx.x * kElimRowsPerGroup + r;
    const int jc = min(j, n - 1);
    double sum = 0.0;
    int s = q;
    for (; s + 3 * kElimSlices < step; s += 4 * kElimSlices) {  // four rows of U in flight
        const double x0 = st.U[(size_t)s * n + jc], x1 = st.U[(size_t)(s + kElimSlices) * n + jc];
        const double x2 = st.U[(size_t)(s + 2 * kElimSlices) * n + jc];
        const double x3 = st.U[(size_t)(s + 3 * kElimSlices) * n + jc];
        sum = fma(x0, u_at[s], sum);
        sum = fma(x1, u_at[s + kElimSlices], sum);
        sum = fma(x2, u_at[s + 2 * kElimSlices], sum);
        sum = fma(x3, u_at[s + 3 * kElimSlices], sum);
    }
    for (; s < step; s += kElimSlices) sum = fma(st.U[(size_t)s * n + jc], u_at[s], sum);
    part[q][r] = sum;
    __syncthreads();
    const bool row_thread = q == 0 && j < n;
    double bj = 0.0, dj = 1.0;
    bool alive = false;
    if (row_thread) {
        double w = W[(size_t)j * n + i] - (part[0][r] + part[1][r] + part[2][r] + part[3][r]);
        if (j == i) w = pivot;
        const double u = w / sqrt(pivot);
        st.U[(size_t)step * n + j] = u;
        bj = st.b[in][j] - w * (bi / pivot);
        dj = st.d[in][j] - u * u;
        alive = st.alive[in][j] != 0 && j != i;
        if (j == i) {
            bj = 0.0;
            dj = 1.0;
        }
        st.b[out][j] = bj;
        st.d[out][j] = dj;
        st.alive[out][j] = alive;
    }
    const int spent = st.spent[in] + (limited && limited[i] != 0 ? 1 : 0);
    const bool allowed = alive && !(limited && limited[j < n ? j : 0] != 0 && spent >= limit);
    group_candidate(bj, dj, allowed, j, row_thread, shared_best, st.best[out] + blockIdx.x);
    if (blockIdx.x == 0 && tid == 0) {
        st.spent[out] = spent;
        removed[step] = i;
    }
}

size_t elimination_workspace(int n, int count) {
    if (n < 0 || count < 0 || count > kElimMaxCount || count > n) return 0;
    return elim_bytes(n, count, (n + kElimRowsPerGroup - 1) / kElimRowsPerGroup);
}

hipError_t launch_elimination(const double* W, const double* b, int n, int count, const int* limited, int limit,
                              long long* removed, void* workspace, hipStream_t st) {
    const int groups = (n + kElimRowsPerGroup - 1) / kElimRowsPerGroup;
    const ElimState s = elim_layout(workspace, n, count, groups);
    hipLaunchKernelGGL(elim_init_kernel, dim3(groups), dim3(kElimThreads), 0, st, W, b, n, limited, limit, s);
    hipError_t e = hipGetLastError();
    for (int step = 0; step < count && e == hipSuccess; ++step) {
        hipLaunchKernelGGL(elim_step_kernel, dim3(groups), dim3(kElimThreads), 0, st, W, n, step, limited, limit, s,
                           removed);
        e = hipGetLastError();
    }
    return e;
}

}  // namespace dsplat
