// Float64 (and float32) Cholesky factorization and the float64 solve of the K x K VarPro /
// Gauss-Newton systems
// (reference: torch.linalg.cholesky_ex and torch.cholesky_solve in dsplat/varpro.py:44-79,
// which run cuSOLVER potrf / potrs; K is about 2,000-2,500 here).
//
// Right-looking blocked factorization with 64 x 64 blocks, two launches per block column:
//   panel:  every workgroup factors the diagonal block in LDS (the same few microseconds of
//           work in each, which saves a launch), then solves its own block row of the panel,
//           X L_kk^T = A_ik, and zeroes the mirrored block of the upper triangle;
//   update: one workgroup per lower trailing block, A_ij -= L_ik L_jk^T on the matrix cores.
// The panel is a chain of 64 dependent pivots, so its cost is latency: the diagonal block is
// factored in four 16-column steps (LAPACK's blocked order at a 16-column inner block): the
// 16 x 16 leaf by one wave with the pivot and column entries passed by lane reads, then the
// rows below the leaf (of the diagonal block and of the workgroup's panel block together)
// by forward substitution, one row per thread, then the trailing columns of both blocks by
// the leaf's 16 columns. Pivot reciprocals scale as LAPACK's potf2 / trsm do (a product by
// 1 / l_jj, here from the hardware reciprocal square root refined by Newton steps). Loops stay
// rolled where an unrolled body would not fit the instruction cache.
// Only the lower triangle of the input is read. info = 1-based index of the first pivot
// that is not positive (0 if none), as LAPACK potrf reports it, in the factor's own precision
// (the float32 form is the first attempt of varpro.py:62-79); the factor is then not
// meaningful, and the launches of the later block columns return at once. The solve runs one
// launch per 64-row block in each direction.
#include "dsplat_internal.h"

namespace dsplat {

constexpr int kCholBlock = 64;
constexpr int kCholLd = kCholBlock + 1;  // LDS row stride of row-major blocks: conflict-free columns
constexpr int kCholLeaf = 16;            // inner block of the panel factorization

// 1 / sqrt(d): the hardware estimate, then Newton steps (two in float64, one in float32),
// y <- y + y (1 - d y^2) / 2.
template <typename T>
__device__ __forceinline__ T rsqrt_newton(T d) {
    if constexpr (sizeof(T) == 8) {
        double y = __builtin_amdgcn_rsq(d);
        y = fma(0.5 * y, fma(-d * y, y, 1.0), y);
        return fma(0.5 * y, fma(-d * y, y, 1.0), y);
    } else {
        const float y = __builtin_amdgcn_rsqf(d);
        return fmaf(0.5f * y, fmaf(-d * y, y, 1.0f), y);
    }
}

// Cholesky of the 16 x 16 leaf at (c, c) of D (row-major, stride kCholLd) by wave 0: lane
// i < 16 keeps row c + i in registers. At step j each lane publishes its entry of column j
// in the wave's column buffer and reads the pivot and the column back (one wave's LDS
// accesses complete in order), so the step is one LDS round trip and the lane's own FMAs.
// Writes the leaf (its lower triangle is the factor) and the pivot reciprocals rd[c + j]; returns the first
// failing step (or -1), uniform over the wave.
template <typename T>
__device__ __forceinline__ int factor_leaf(T* D, T* rd, T* col, int c, int lane) {
    // entries above the diagonal are carried along unread: the leaf's upper triangle is never
    // used, and the diagonal block's final store writes zeros there
    const int i = lane & (kCholLeaf - 1);
    T x[kCholLeaf];
#pragma unroll
    for (int m = 0; m < kCholLeaf; ++m) x[m] = D[(c + i) * kCholLd + c + m];
    int first = -1;
#pragma unroll
    for (int j = 0; j < kCholLeaf; ++j) {
        if (lane < kCholLeaf) col[lane] = x[j];  // a_kj, k = lane (entries k < j unused)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        T a[kCholLeaf];
#pragma unroll
        for (int k = j; k < kCholLeaf; ++k) a[k] = col[k];
        const T d = a[j];
        first = (first < 0 && !(d > T(0))) ? j : first;
        const T r = rsqrt_newton(d);
        const T l = x[j] * r;  // l_ij; on lane j x[j] = d, so l_jj = d / sqrt(d)
        const T t = l * r;
#pragma unroll
        for (int k = j + 1; k < kCholLeaf; ++k) x[k] = fma(-t, a[k], x[k]);  // a_ik -= l_ij l_kj
        x[j] = l;
        if (lane == 0) rd[c + j] = r;
        // the next step's write follows this step's reads in the wave's LDS order
    }
    if (lane < kCholLeaf) {
#pragma unroll
        for (int m = 0; m < kCholLeaf; ++m) D[(c + i) * kCholLd + c + m] = x[m];
    }
    return __builtin_amdgcn_readfirstlane(first);
}

// Row y (16 values at columns c..c+15 of a row-major LDS block) solved against the leaf:
// y L_leaf^T = b by forward substitution.
template <typename T>
__device__ __forceinline__ void solve_leaf_row(T* y, const T* D, const T* rd, int c) {
    T v[kCholLeaf];
#pragma unroll
    for (int m = 0; m < kCholLeaf; ++m) v[m] = y[c + m];
#pragma unroll
    for (int j = 0; j < kCholLeaf; ++j) {
        T s = v[j];
#pragma unroll
        for (int m = 0; m < j; ++m) s = fma(-v[m], D[(c + j) * kCholLd + c + m], s);
        v[j] = s * rd[c + j];
    }
#pragma unroll
    for (int m = 0; m < kCholLeaf; ++m) y[c + m] = v[m];
}

// Factor the 64 x 64 diagonal block D (lower triangle; rows >= w padded with the identity)
// and, when B is given, solve its h rows: B <- B L^-T. rd receives the pivot reciprocals.
// Returns the first failing pivot index (or -1) on every thread.
template <typename T>
__device__ int factor_panel(T* D, T* B, T* rd, int w) {
    __shared__ int fail;
    __shared__ T col[kCholLeaf];
    const int tid = threadIdx.x, lane = tid & 63;
    if (tid == 0) fail = -1;
    for (int c = 0; c < kCholBlock; c += kCholLeaf) {
        if (tid < 64) {
            const int f = factor_leaf(D, rd, col, c, lane);
            if (tid == 0 && f >= 0 && fail < 0 && c + f < w) fail = c + f;
        }
        __syncthreads();
        // rows below the leaf: D rows c + 16 .. 63 (threads 64..), B rows 0..63 (threads 0..63)
        const int below = kCholBlock - c - kCholLeaf;
        if (tid < 64 && B) {
            solve_leaf_row(B + tid * kCholLd, D, rd, c);
        } else if (tid >= 64 && tid < 64 + below) {
            solve_leaf_row(D + (c + kCholLeaf + tid - 64) * kCholLd, D, rd, c);
        }
        __syncthreads();
        if (below > 0) {
            // trailing columns s in [c + 16, 64) of D rows c + 16.. and of B rows:
            // row index rho = tid % 128 (B rows, then D rows), columns of parity tid / 128
            const int rho = tid & 127, par = tid >> 7;
            T* row = nullptr;
            if (rho < 64) {
                if (B) row = B + rho * kCholLd;
            } else if (rho - 64 < below) {
                row = D + (c + kCholLeaf + rho - 64) * kCholLd;
            }
            if (row) {
                T u[kCholLeaf];
#pragma unroll
                for (int m = 0; m < kCholLeaf; ++m) u[m] = row[c + m];
                // two columns per pass: two independent FMA chains
                for (int s = c + kCholLeaf + par; s < kCholBlock; s += 4) {
                    const bool two = s + 2 < kCholBlock;
                    const T* l0 = D + s * kCholLd + c;
                    const T* l1 = D + (two ? s + 2 : s) * kCholLd + c;
                    T a0 = row[s], a1 = two ? row[s + 2] : T(0);
#pragma unroll
                    for (int m = 0; m < kCholLeaf; ++m) {
                        a0 = fma(-u[m], l0[m], a0);
                        a1 = fma(-u[m], l1[m], a1);
                    }
                    row[s] = a0;
                    if (two) row[s + 2] = a1;
                }
            }
        }
        __syncthreads();
    }
    return fail;
}

template <typename T>
__global__ __launch_bounds__(256) void chol_panel_kernel(T* __restrict__ A, int n, int lda, int kb,
                                                        int* __restrict__ info) {
    __shared__ T D[kCholBlock * kCholLd];
    __shared__ T Bs[kCholBlock * kCholLd];
    __shared__ T rd[kCholBlock];
    // an earlier block column failed: the factor is not meaningful, info is final
    if (*info <= n) return;
    const int tid = threadIdx.x;
    const int c0 = kb * kCholBlock, w = min(kCholBlock, n - c0);
    const int ib = kb + blockIdx.x;
    const int r0 = ib * kCholBlock, h = min(kCholBlock, n - r0);
    const bool diag = blockIdx.x == 0;
    // all loads in flight before the first LDS store
    constexpr int kPer = kCholBlock * kCholBlock / 256;
    T vd[kPer], vb[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int e = tid + 256 * q, i = e >> 6, k = e & 63;
        vd[q] = (i < w && k <= i) ? A[(size_t)(c0 + i) * lda + c0 + k] : T(i >= w && k == i ? 1 : 0);
        vb[q] = (!diag && i < h && k < w) ? A[(size_t)(r0 + i) * lda + c0 + k] : T(0);
    }
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int e = tid + 256 * q, i = e >> 6, k = e & 63;
        D[i * kCholLd + k] = vd[q];
        Bs[i * kCholLd + k] = vb[q];
    }
    __syncthreads();
    const int fail = factor_panel(D, diag ? nullptr : Bs, rd, w);
    if (diag) {
        if (tid == 0 && fail >= 0) atomicMin(info, c0 + fail + 1);
        for (int e = tid; e < w * kCholBlock; e += 256) {
            const int i = e >> 6, k = e & 63;
            if (k < w) A[(size_t)(c0 + i) * lda + c0 + k] = k <= i ? D[i * kCholLd + k] : T(0);
        }
        return;
    }
    for (int e = tid; e < h * kCholBlock; e += 256) {
        const int i = e >> 6, k = e & 63;
        if (k < w) A[(size_t)(r0 + i) * lda + c0 + k] = Bs[i * kCholLd + k];
    }
    // the mirrored block of the upper triangle: rows c0.., columns r0.. (coalesced along i)
    for (int e = tid; e < w * kCholBlock; e += 256) {
        const int k = e >> 6, i = e & 63;
        if (i < h) A[(size_t)(c0 + k) * lda + r0 + i] = T(0);
    }
}

// A_ij -= L_ik L_jk^T for the trailing lower blocks kb < jb <= ib on the matrix cores
// (v_mfma_f64_16x16x4f64, or v_mfma_f32_16x16x4f32 for float32): each wave computes a 32 x 32
// quarter of the 64 x 64 tile as 2 x 2 products of 16 x 16, K = the panel's 64 columns in
// steps of 4. Operand lane l reads row l % 16, column l / 16 of a step: A[i][k] and
// B[k][j] = L_j[j][k] alike; result register r holds column l % 16 of row 4 (l / 16) + r
// (float32) or row l / 16 + 4 r (float64). The panel
// blocks are staged row-major with a row stride of 68 words (banks of a 16-row x 4-column read
// spread), a diagonal tile stages one block.
template <typename T>
__global__ __launch_bounds__(256) void chol_update_kernel(T* __restrict__ A, int n, int lda, int kb,
                                                         const int* __restrict__ info) {
    constexpr int LD = kCholBlock + 4;
    __shared__ T Li[kCholBlock * LD];
    __shared__ T Lj[kCholBlock * LD];
    if (*info <= n) return;  // a pivot failed: nothing left to compute
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
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
    const bool same = ti == tj;
    // all loads in flight before the first LDS store
    constexpr int kPer = kCholBlock * kCholBlock / 256;
    T vi[kPer], vj[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int e = tid + 256 * q, r = e >> 6, p = e & 63;
        vi[q] = (r < hi && p < w) ? A[(size_t)(ri + r) * lda + c0 + p] : T(0);
        vj[q] = (!same && r < hj && p < w) ? A[(size_t)(rj + r) * lda + c0 + p] : T(0);
    }
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int e = tid + 256 * q, r = e >> 6, p = e & 63;
        Li[r * LD + p] = vi[q];
        Lj[r * LD + p] = vj[q];
    }
    __syncthreads();
    const T* Lb = same ? Li : Lj;
    using V4 = T __attribute__((ext_vector_type(4)));
    const int row0 = 32 * (wave >> 1), col0 = 32 * (wave & 1);
    const int i16 = lane & 15, k4 = lane >> 4;
    V4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = V4{T(0), T(0), T(0), T(0)};
    // columns p >= w are staged as zeros: all 16 steps run
#pragma unroll 4
    for (int s4 = 0; s4 < kCholBlock / 4; ++s4) {
        const int k = 4 * s4 + k4;
        T x[2], y[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) x[a] = Li[(row0 + 16 * a + i16) * LD + k];
#pragma unroll
        for (int b = 0; b < 2; ++b) y[b] = Lb[(col0 + 16 * b + i16) * LD + k];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                if constexpr (sizeof(T) == 8)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[a], y[b], acc[a][b], 0, 0, 0);
                else
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[a], y[b], acc[a][b], 0, 0, 0);
            }
    }
    // the tile's current values are all loaded before the first store (a store to A would
    // otherwise order every later load of A behind it)
    T old[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                // result rows: 4 (l / 16) + r for the float32 product, l / 16 + 4 r for float64
                const int rr = sizeof(T) == 8 ? k4 + 4 * r : 4 * k4 + r;
                const int row = row0 + 16 * a + rr, cc = col0 + 16 * b + i16;
                old[a][b][r] = (row < hi && cc < hj) ? A[(size_t)(ri + row) * lda + rj + cc] : T(0);
            }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = sizeof(T) == 8 ? k4 + 4 * r : 4 * k4 + r;
                const int row = row0 + 16 * a + rr, cc = col0 + 16 * b + i16;
                if (row < hi && cc < hj) A[(size_t)(ri + row) * lda + rj + cc] = old[a][b][r] - acc[a][b][r];
            }
}

// value of v on lane k (k uniform)
__device__ __forceinline__ double lane_value(double v, int k) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)b, k);
    const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(b >> 32), k);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Solve L L^T X = B in place for nrhs columns of B (column j at B + j * ldb): forward
// L Y = B into the workspace W, then backward L^T X = Y into B, one launch per 64-row block in
// each direction. Block b's launch: every workgroup stages the diagonal block and solves its
// triangle for the block's rows (one wave: lane = row; at step m the value of unknown m reaches
// every lane by a lane read, and lanes subtract their entry of column m times it; the staged
// triangle holds zeros off its strict part, so no lane needs a mask); workgroup 0 stores the
// solved rows in the other array and workgroup g >= 1 subtracts their product with its 64-row
// slice of the off-diagonal part from the rows still to be solved. Rows a launch reads are
// written by none of its workgroups. The slice is staged while the triangle is solved.
constexpr int kTriMaxRhs = 8;

template <bool BACKWARD>
__global__ __launch_bounds__(256) void tri_block_kernel(const double* __restrict__ L, int n, int lda,
                                                        double* __restrict__ R, int ldr, double* __restrict__ X,
                                                        int ldx, int nrhs, int blk) {
    __shared__ double T[kCholBlock * kCholLd];  // strict triangle, row = unknown's lane
    __shared__ double S[kCholBlock * kCholLd];  // the workgroup's off-diagonal slice
    __shared__ double rinv[kCholBlock];
    __shared__ double Y[kTriMaxRhs][kCholBlock];
    __shared__ double part[4][kCholBlock];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r0 = blk * kCholBlock, h = min(kCholBlock, n - r0);
    const int g = blockIdx.x;
    // slice rows: below the block (forward) or above it (backward)
    const int c0 = BACKWARD ? (g - 1) * kCholBlock : r0 + g * kCholBlock;
    const int hs = g == 0 ? 0 : min(kCholBlock, BACKWARD ? r0 - c0 : n - c0);
    constexpr int kPer = kCholBlock * kCholBlock / 256;
    double vt[kPer], vs[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int e = tid + 256 * q, i = e >> 6, m = e & 63;
        // forward: T[i][m] = L[r0 + i][r0 + m] (m < i); backward: T[i][m] = L[r0 + m][r0 + i] (m > i)
        const bool strict = BACKWARD ? (m > i && m < h) : (m < i && i < h);
        vt[q] = strict ? (BACKWARD ? L[(size_t)(r0 + m) * lda + r0 + i] : L[(size_t)(r0 + i) * lda + r0 + m]) : 0.0;
        // forward slice S[i][m] = L[c0 + i][r0 + m]; backward S[m][c] = L[r0 + m][c0 + c] (i = m, m = c)
        vs[q] = 0.0;
        if (BACKWARD) {
            if (i < h && m < hs) vs[q] = L[(size_t)(r0 + i) * lda + c0 + m];
        } else {
            if (i < hs && m < h) vs[q] = L[(size_t)(c0 + i) * lda + r0 + m];
        }
    }
    if (tid < kCholBlock) rinv[tid] = tid < h ? 1.0 / L[(size_t)(r0 + tid) * lda + r0 + tid] : 0.0;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int e = tid + 256 * q, i = e >> 6, m = e & 63;
        T[i * kCholLd + m] = vt[q];
        S[i * kCholLd + m] = vs[q];
    }
    __syncthreads();
    if (wave == 0) {
        const double ri = rinv[lane];
        for (int j = 0; j < nrhs; ++j) {
            double sv = lane < h ? R[(size_t)j * ldr + r0 + lane] : 0.0;
            if (BACKWARD) {
#pragma unroll
                for (int m = kCholBlock - 1; m >= 0; --m) {
                    const double um = lane_value(sv * ri, m);
                    sv = fma(-T[lane * kCholLd + m], um, sv);
                }
            } else {
#pragma unroll
                for (int m = 0; m < kCholBlock; ++m) {
                    const double um = lane_value(sv * ri, m);
                    sv = fma(-T[lane * kCholLd + m], um, sv);
                }
            }
            Y[j][lane] = sv * ri;  // lane i's value is final once step i has passed
        }
    }
    __syncthreads();
    if (g == 0) {
        for (int e = tid; e < nrhs * kCholBlock; e += 256) {
            const int j = e >> 6, i = e & 63;
            if (i < h) X[(size_t)j * ldx + r0 + i] = Y[j][i];
        }
        return;
    }
    // R[c] -= sum_m S(c, m) Y[m] over the slice's rows c; thread (c = lane, quarter = wave)
    for (int j = 0; j < nrhs; ++j) {
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < kCholBlock / 4; ++k) {
            const int m = wave + 4 * k;
            const double sv = BACKWARD ? S[m * kCholLd + lane] : S[lane * kCholLd + m];
            acc = fma(sv, Y[j][m], acc);
        }
        part[wave][lane] = acc;
        __syncthreads();
        if (tid < hs) {
            const double sub = (part[0][tid] + part[1][tid]) + (part[2][tid] + part[3][tid]);
            R[(size_t)j * ldr + c0 + tid] -= sub;
        }
        __syncthreads();
    }
}

// One power-iteration product for this workgroup's rows: yout = h x with x = yin /
// max(|yin|, tiny) (x = yin on the first step, the start vector). Every workgroup forms |yin|
// in the same fixed order, so all normalize identically; one wave per row pair, lanes over
// the columns.
constexpr int kPowerRows = 8;
__global__ __launch_bounds__(256) void power_step_kernel(const double* __restrict__ h, int n, int ld,
                                                         const double* __restrict__ yin, int normalize,
                                                         double* __restrict__ yout) {
    extern __shared__ double xs[];  // [n]
    __shared__ double red[256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double top = 1.0;
    if (normalize) {
        double sq = 0.0;
        for (int c = tid; c < n; c += 256) sq = fma(yin[c], yin[c], sq);
        red[tid] = sq;
        __syncthreads();
        for (int w = 128; w >= 1; w >>= 1) {
            if (tid < w) red[tid] += red[tid + w];
            __syncthreads();
        }
        top = fmax(sqrt(red[0]), 2.2250738585072014e-308);
    }
    for (int c = tid; c < n; c += 256) xs[c] = normalize ? yin[c] / top : yin[c];
    __syncthreads();
    // the wave's two rows together, four columns per lane and row in flight per pass
    const int r0 = blockIdx.x * kPowerRows + 2 * wave;
    if (r0 >= n) return;  // uniform over the wave; no barrier follows
    const bool two = r0 + 1 < n;
    const double* row0 = h + (size_t)r0 * ld;
    const double* row1 = h + (size_t)(two ? r0 + 1 : r0) * ld;
    double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;
    int c = lane;
    for (; c + 192 < n; c += 256) {
        const double x0 = xs[c], x1 = xs[c + 64], x2 = xs[c + 128], x3 = xs[c + 192];
        const double p0 = row0[c], p1 = row0[c + 64], p2 = row0[c + 128], p3 = row0[c + 192];
        const double q0 = row1[c], q1 = row1[c + 64], q2 = row1[c + 128], q3 = row1[c + 192];
        a0 = fma(p0, x0, a0);
        a1 = fma(p1, x1, a1);
        a0 = fma(p2, x2, a0);
        a1 = fma(p3, x3, a1);
        b0 = fma(q0, x0, b0);
        b1 = fma(q1, x1, b1);
        b0 = fma(q2, x2, b0);
        b1 = fma(q3, x3, b1);
    }
    for (; c < n; c += 64) {
        a0 = fma(row0[c], xs[c], a0);
        b0 = fma(row1[c], xs[c], b0);
    }
    double ya = a0 + a1, yb = b0 + b1;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        ya += __shfl_xor(ya, o, 64);
        yb += __shfl_xor(yb, o, 64);
    }
    if (lane == 0) {
        yout[r0] = ya;
        if (two) yout[r0 + 1] = yb;
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

hipError_t launch_cholesky_solve(const double* L, int n, int lda, double* B, int nrhs, int ldb, double* W,
                                 hipStream_t st) {
    const int nblk = (n + kCholBlock - 1) / kCholBlock;
    for (int j0 = 0; j0 < nrhs; j0 += kTriMaxRhs) {
        const int m = min(kTriMaxRhs, nrhs - j0);
        double* Bj = B + (size_t)j0 * ldb;
        for (int b = 0; b < nblk; ++b) {
            hipLaunchKernelGGL(tri_block_kernel<false>, dim3(nblk - b), dim3(256), 0, st, L, n, lda, Bj, ldb, W, n, m,
                               b);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        for (int b = nblk - 1; b >= 0; --b) {
            hipLaunchKernelGGL(tri_block_kernel<true>, dim3(1 + b), dim3(256), 0, st, L, n, lda, W, n, Bj, ldb, m, b);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

size_t cholesky_solve_workspace(int n, int nrhs) {
    return sizeof(double) * (size_t)n * (size_t)min(nrhs, kTriMaxRhs);
}

hipError_t launch_power_iteration(const double* h, int n, int ld, int steps, const double* start, double* out,
                                  double* work, hipStream_t st) {
    const size_t smem = sizeof(double) * (size_t)n;
    hipError_t e = allow_lds((const void*)power_step_kernel, smem + sizeof(double) * 256);
    if (e != hipSuccess) return e;
    const dim3 grid((n + kPowerRows - 1) / kPowerRows);
    for (int i = 0; i < steps; ++i) {
        const double* in = i == 0 ? start : work + (size_t)((i - 1) & 1) * n;
        double* o = i == steps - 1 ? out : work + (size_t)(i & 1) * n;
        hipLaunchKernelGGL(power_step_kernel, grid, dim3(256), smem, st, h, n, ld, in, (int)(i > 0), o);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace dsplat
