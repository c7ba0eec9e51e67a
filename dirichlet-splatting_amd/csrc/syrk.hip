// Symmetric rank-k update on MFMA: upper(out) += rows^T rows (reference: the cuBLAS SSYRK
// call of dsplat/kernels/cublas.py:31-44; the lower triangle is left untouched).
//
// rows [m, n] row-major, out [n, n] row-major. A workgroup (4 waves) owns a 128 x 128
// tile of the upper triangle and a range of rows (split-K); with col0 >= 0 the tiles cover
// instead the rectangle out [n, n - col0] += rows^T rows[:, col0:] (the cross products of new
// columns against all of them, the pooled statistics' GEMM); each wave accumulates a
// 64 x 64 block as 2 x 2 MFMA blocks of 32 x 32. Rows are staged through LDS in chunks
// of 32 (MODE 0) or 16 (MODES 1, 2: 48 KB of LDS, three workgroups per CU), column-major within
// a panel (k contiguous) so each lane reads its K values with 128-bit loads, double-buffered
// with the next chunk's global loads in flight.
//   MODE 0: FP32 operands, v_mfma_f32_32x32x2f32 (the lane's two K rows are permuted to
//           h * 16 + kk; the sum over K does not depend on its order).
//   MODE 1: BF16 hi + lo split of each value, hi hi + hi lo + lo hi on
//           v_mfma_f32_32x32x16_bf16 (relative product error about 2^-16).
//   MODE 2: FP16 hi + lo split with each column scaled by 2^e_c into [2^13, 2^14) from its
//           largest magnitude (relative product error about 2^-22).
// Partials [split][tile][128][128] are summed in float64 in a fixed order, unscaled, and
// added to the upper triangle of out by syrk_reduce: deterministic, no atomics on out.
// Lane layouts (gfx950): 32x32x2f32 A[i][k] -> lane (i + 32 k); 32x32x16 A[i][8 h + e] ->
// lane (i + 32 h) element e; B likewise with j; C[8 (r / 4) + 4 (lane / 32) + r % 4][lane % 32].
#include "dsplat_internal.h"

namespace dsplat {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

constexpr int kTile = 128;
constexpr int kChunk = 32;            // rows per LDS chunk, MODE 0; split rows are a multiple of it
constexpr int kChunkH = 16;           // rows per LDS chunk, MODES 1 and 2 (48 KB of LDS: 3 workgroups per CU)
constexpr int kStrideF = kChunk + 4;  // floats per LDS column (MODE 0)
constexpr int kStrideH = kChunkH + 8; // 16-bit values per LDS column (MODES 1, 2): 48 B, so the 16-byte
                                      // accesses of 16 consecutive columns fall on distinct bank groups
constexpr int kMinSplitRows = 512;
constexpr int kMaxSplitRows = 16384;  // rows per float32 partial; the reduction beyond is float64

template <int MODE>
struct Half16;
template <>
struct Half16<1> {
    typedef __bf16 scalar;
    typedef bf16x8 vec;
};
template <>
struct Half16<2> {
    typedef _Float16 scalar;
    typedef f16x8 vec;
};

__device__ __forceinline__ f32x16 mma16(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 mma16(f16x8 a, f16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void syrk_tile_of(int t, int nb, int* bi, int* bj) {
    // t-th tile of the upper triangle, row by row: row bi holds tiles bi..nb-1
    int row = 0;
    while (t >= nb - row) {
        t -= nb - row;
        ++row;
    }
    *bi = row;
    *bj = row + t;
}

// First output row and column of tile t: the upper triangle of rows^T rows (col0 < 0), or
// the full rectangle rows^T rows[:, col0:] (columns col0..n-1) row-major over tiles.
__device__ __forceinline__ void tile_origin(int t, int n, int col0, int* ci, int* cj) {
    const int nb = (n + kTile - 1) / kTile;
    if (col0 < 0) {
        int bi, bj;
        syrk_tile_of(t, nb, &bi, &bj);
        *ci = bi * kTile;
        *cj = bj * kTile;
    } else {
        const int nbj = (n - col0 + kTile - 1) / kTile;
        *ci = (t / nbj) * kTile;
        *cj = col0 + (t % nbj) * kTile;
    }
}

// Exponent that scales a column whose largest magnitude is `top` into [2^13, 2^14).
__device__ __forceinline__ int column_exponent(float top) {
    if (!(top > 0.0f) || !isfinite(top)) return 0;
    int p;
    frexpf(top, &p);
    return min(100, max(-100, 14 - p));
}

// colmax[c] = max |rows[k][c]| as float bits (non-negative floats order like unsigned ints).
__global__ __launch_bounds__(256) void syrk_colmax(const float* __restrict__ rows, int m, int n, int rows_per,
                                                   unsigned* __restrict__ colmax) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const int k0 = blockIdx.y * rows_per, k1 = min(m, k0 + rows_per);
    float top = 0.0f;
    for (int k = k0; k < k1; ++k) top = fmaxf(top, fabsf(rows[(size_t)k * n + c]));
    atomicMax(colmax + c, __float_as_uint(top));
}

// Global -> registers: thread t holds column (t & 127) of a panel, rows (CH / 2) (t >> 7) + 0..CH/2-1.
template <int CH>
struct Staged {
    float v[2][CH / 2];  // [panel I / J][row]
};

template <int CH>
__device__ __forceinline__ void stage_load(Staged<CH>& s, const float* __restrict__ rows, int n, int k0, int k_end,
                                           int ci, int cj, bool two, int tid) {
    constexpr int R = CH / 2;
    const int c = tid & (kTile - 1), rg = tid >> 7;
    if (k0 + CH <= k_end && ci + kTile <= n && (!two || cj + kTile <= n)) {
        // a full chunk of full panels (the common case): plain loads, no per-load predicate;
        // a diagonal tile (one panel) loads it twice and never reads the second copy
        const float* pi = rows + (size_t)(k0 + rg * R) * n + ci + c;
        const float* pj = rows + (size_t)(k0 + rg * R) * n + (two ? cj : ci) + c;
#pragma unroll
        for (int e = 0; e < R; ++e) {
            s.v[0][e] = pi[(size_t)e * n];
            s.v[1][e] = pj[(size_t)e * n];
        }
        return;
    }
#pragma unroll
    for (int e = 0; e < R; ++e) {
        const int k = k0 + rg * R + e;
        const bool in = k < k_end;
        s.v[0][e] = in && ci + c < n ? rows[(size_t)k * n + ci + c] : 0.0f;
        s.v[1][e] = two && in && cj + c < n ? rows[(size_t)k * n + cj + c] : 0.0f;
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void syrk_kernel(const float* __restrict__ rows, int m, int n, int col0,
                                                   int per_split, const unsigned* __restrict__ colmax,
                                                   float* __restrict__ partial) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    // the tiles of one row range run on one XCD: the panels they share come through one L2
    const int lin = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int tile = lin % gridDim.x, split = lin / gridDim.x;
    int ci, cj;
    tile_origin(tile, n, col0, &ci, &cj);
    const bool two = ci != cj;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wi = wave >> 1, wj = wave & 1, h = lane >> 5, l32 = lane & 31;
    const int k_begin = split * per_split, k_end = min(m, k_begin + per_split);
    const int c = tid & (kTile - 1), rg = tid >> 7;

    f32x16 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[x][y][r] = 0.0f;

    // per-thread column scales of the staged columns (MODE 2)
    float scale_i = 1.0f, scale_j = 1.0f;
    if constexpr (MODE == 2) {
        if (ci + c < n) scale_i = ldexpf(1.0f, column_exponent(__uint_as_float(colmax[ci + c])));
        if (two && cj + c < n) scale_j = ldexpf(1.0f, column_exponent(__uint_as_float(colmax[cj + c])));
    }

    constexpr int CH = MODE == 0 ? kChunk : kChunkH;
    Staged<CH> st;
    int buf = 0;
    if (k_begin < k_end) stage_load<CH>(st, rows, n, k_begin, k_end, ci, cj, two, tid);

    if constexpr (MODE == 0) {
        float* X = reinterpret_cast<float*>(lds);  // [2 buf][2 panel][128][kStrideF]
        auto col_ptr = [&](int b, int p, int colx) { return X + ((b * 2 + p) * kTile + colx) * kStrideF; };
        auto store = [&](int b) {
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                float* dst = col_ptr(b, p, c) + rg * 16;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    *reinterpret_cast<f32x4*>(dst + 4 * q) =
                        f32x4{st.v[p][4 * q], st.v[p][4 * q + 1], st.v[p][4 * q + 2], st.v[p][4 * q + 3]};
            }
        };
        if (k_begin < k_end) store(0);
        __syncthreads();
        for (int k0 = k_begin; k0 < k_end; k0 += CH) {
            const bool more = k0 + CH < k_end;
            if (more) stage_load<CH>(st, rows, n, k0 + CH, k_end, ci, cj, two, tid);
            const int pj = two ? 1 : 0;
            float a[2][16], b[2][16];
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                const float* src = col_ptr(buf, 0, wi * 64 + x * 32 + l32) + h * 16;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const f32x4 v = *reinterpret_cast<const f32x4*>(src + 4 * q);
                    a[x][4 * q] = v[0], a[x][4 * q + 1] = v[1], a[x][4 * q + 2] = v[2], a[x][4 * q + 3] = v[3];
                }
            }
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                const float* src = col_ptr(buf, pj, wj * 64 + y * 32 + l32) + h * 16;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const f32x4 v = *reinterpret_cast<const f32x4*>(src + 4 * q);
                    b[y][4 * q] = v[0], b[y][4 * q + 1] = v[1], b[y][4 * q + 2] = v[2], b[y][4 * q + 3] = v[3];
                }
            }
#pragma unroll
            for (int kk = 0; kk < 16; ++kk)
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int y = 0; y < 2; ++y)
                        acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[x][kk], b[y][kk], acc[x][y], 0, 0, 0);
            if (more) store(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
    } else {
        typedef typename Half16<MODE>::scalar h16;
        typedef typename Half16<MODE>::vec v8;
        h16* X = reinterpret_cast<h16*>(lds);  // [2 buf][2 part hi/lo][2 panel][128][kStrideH]
        auto col_ptr = [&](int b, int part, int p, int colx) {
            return X + (((b * 2 + part) * 2 + p) * kTile + colx) * kStrideH;
        };
        auto store = [&](int b) {
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const float sc = p == 0 ? scale_i : scale_j;
                constexpr int NV = CH / 16;  // 8-element vectors per thread and panel
                v8 hi[NV], lo[NV];
                if constexpr (MODE == 2) {
#pragma unroll
                    for (int e = 0; e < CH / 2; e += 2) {
                        f16x2 h2, l2;
                        split_f16_pair<false>(st.v[p][e] * sc, st.v[p][e + 1] * sc, h2, l2);
                        hi[e >> 3][e & 7] = h2.x, hi[e >> 3][(e & 7) + 1] = h2.y;
                        lo[e >> 3][e & 7] = l2.x, lo[e >> 3][(e & 7) + 1] = l2.y;
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < CH / 2; ++e) {
                        const float x = st.v[p][e];
                        const h16 xh = (h16)x;
                        hi[e >> 3][e & 7] = xh;
                        lo[e >> 3][e & 7] = (h16)(x - (float)xh);
                    }
                }
                v8* dh = reinterpret_cast<v8*>(col_ptr(b, 0, p, c) + rg * (CH / 2));
                v8* dl = reinterpret_cast<v8*>(col_ptr(b, 1, p, c) + rg * (CH / 2));
#pragma unroll
                for (int q = 0; q < NV; ++q) {
                    dh[q] = hi[q];
                    dl[q] = lo[q];
                }
            }
        };
        if (k_begin < k_end) store(0);
        __syncthreads();
        for (int k0 = k_begin; k0 < k_end; k0 += CH) {
            const bool more = k0 + CH < k_end;
            if (more) stage_load<CH>(st, rows, n, k0 + CH, k_end, ci, cj, two, tid);
            const int pj = two ? 1 : 0;
#pragma unroll
            for (int s = 0; s < CH / 16; ++s) {
                v8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
                for (int x = 0; x < 2; ++x) {
                    const int colx = wi * 64 + x * 32 + l32;
                    ah[x] = *reinterpret_cast<const v8*>(col_ptr(buf, 0, 0, colx) + s * 16 + 8 * h);
                    al[x] = *reinterpret_cast<const v8*>(col_ptr(buf, 1, 0, colx) + s * 16 + 8 * h);
                }
#pragma unroll
                for (int y = 0; y < 2; ++y) {
                    const int colx = wj * 64 + y * 32 + l32;
                    bh[y] = *reinterpret_cast<const v8*>(col_ptr(buf, 0, pj, colx) + s * 16 + 8 * h);
                    bl[y] = *reinterpret_cast<const v8*>(col_ptr(buf, 1, pj, colx) + s * 16 + 8 * h);
                }
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int y = 0; y < 2; ++y) {
                        f32x16 t = acc[x][y];
                        t = mma16(ah[x], bh[y], t);
                        t = mma16(ah[x], bl[y], t);
                        t = mma16(al[x], bh[y], t);
                        acc[x][y] = t;
                    }
            }
            if (more) store(buf ^ 1);
            __syncthreads();
            buf ^= 1;
        }
    }

    float* dst = partial + ((size_t)split * gridDim.x + tile) * kTile * kTile;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int col = wj * 64 + y * 32 + l32;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = wi * 64 + x * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
                dst[row * kTile + col] = acc[x][y][r];
            }
        }
}

// out[gi][gj] += sum over splits (fixed order, float64), unscaled by 2^-(e_i + e_j) in
// MODE 2: gi <= gj of out [n, n] for the triangle, every entry of out [n, n - col0]
// (column gj - col0) for the rectangle.
__global__ __launch_bounds__(256) void syrk_reduce(const float* __restrict__ partial, int n, int col0, int tiles,
                                                   int splits, const unsigned* __restrict__ colmax,
                                                   float* __restrict__ out) {
    int ci, cj;
    tile_origin(blockIdx.x, n, col0, &ci, &cj);
    const int width = col0 < 0 ? n : n - col0, shift = col0 < 0 ? 0 : col0;
    for (int e = blockIdx.y * 256 + threadIdx.x; e < kTile * kTile; e += gridDim.y * 256) {
        const int row = e / kTile, col = e % kTile;
        const int gi = ci + row, gj = cj + col;
        if (gi >= n || gj >= n || (col0 < 0 && gi > gj)) continue;
        double sum = 0.0;
        for (int s = 0; s < splits; ++s) sum += partial[((size_t)s * tiles + blockIdx.x) * kTile * kTile + e];
        if (colmax)
            sum = ldexp(sum, -column_exponent(__uint_as_float(colmax[gi])) -
                                 column_exponent(__uint_as_float(colmax[gj])));
        out[(size_t)gi * width + (gj - shift)] += (float)sum;
    }
}

struct SyrkShape {
    int tiles, splits, per_split;
};

static SyrkShape syrk_shape(int m, int n, int col0) {
    SyrkShape s;
    const int nb = (n + kTile - 1) / kTile;
    s.tiles = col0 < 0 ? nb * (nb + 1) / 2 : nb * ((n - col0 + kTile - 1) / kTile);
    int splits = (1024 + s.tiles - 1) / s.tiles;
    const int most = (m + kMinSplitRows - 1) / kMinSplitRows;
    if (splits > most) splits = most;
    // float32 accumulators sum at most kMaxSplitRows rows: long blocks get more ranges
    const int fold = (int)(((long long)m + kMaxSplitRows - 1) / kMaxSplitRows);
    if (splits < fold) splits = fold;
    if (splits > 65535) splits = 65535;  // grid y
    if (splits < 1) splits = 1;
    s.per_split = (m + splits - 1) / splits;
    s.per_split = (s.per_split + kChunk - 1) / kChunk * kChunk;
    s.splits = (m + s.per_split - 1) / s.per_split;
    if (s.splits < 1) s.splits = 1;
    return s;
}

size_t syrk_workspace(int m, int n, int col0, int mode) {
    if (m <= 0 || n <= 0 || col0 >= n) return 0;
    const SyrkShape s = syrk_shape(m, n, col0);
    return sizeof(float) * (size_t)s.splits * s.tiles * kTile * kTile + (mode == 2 ? 256 + sizeof(unsigned) * n : 0);
}

static size_t syrk_lds(int mode) {
    return mode == 0 ? sizeof(float) * 2 * 2 * kTile * kStrideF : sizeof(short) * 2 * 2 * 2 * kTile * kStrideH;
}

template <int MODE>
static hipError_t launch_syrk_mode(const float* rows, int m, int n, int col0, const SyrkShape& s,
                                   const unsigned* colmax, float* partial, hipStream_t st) {
    const size_t smem = syrk_lds(MODE);
    hipError_t e = allow_lds((const void*)syrk_kernel<MODE>, smem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((syrk_kernel<MODE>), dim3(s.tiles, s.splits), dim3(256), smem, st, rows, m, n, col0,
                       s.per_split, colmax, partial);
    return hipGetLastError();
}

hipError_t launch_syrk(const float* rows, int m, int n, int col0, int mode, const unsigned* column_max, float* out,
                       void* workspace, hipStream_t st) {
    if (m <= 0 || n <= 0 || col0 >= n) return hipSuccess;
    const SyrkShape s = syrk_shape(m, n, col0);
    float* partial = static_cast<float*>(workspace);
    const unsigned* colmax = mode == 2 ? column_max : nullptr;
    hipError_t e;
    if (mode == 2 && !column_max) {
        unsigned* colmax_ws = reinterpret_cast<unsigned*>(partial + (size_t)s.splits * s.tiles * kTile * kTile);
        colmax = colmax_ws;
        if ((e = fill_zero(reinterpret_cast<float*>(colmax_ws), n, st)) != hipSuccess) return e;
        const int rows_per = (m + 4095) / 4096 > 256 ? (m + 4095) / 4096 : 256;
        hipLaunchKernelGGL(syrk_colmax, dim3((n + 255) / 256, (m + rows_per - 1) / rows_per), dim3(256), 0, st,
                           rows, m, n, rows_per, colmax_ws);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    switch (mode) {
        case 0: e = launch_syrk_mode<0>(rows, m, n, col0, s, colmax, partial, st); break;
        case 1: e = launch_syrk_mode<1>(rows, m, n, col0, s, colmax, partial, st); break;
        case 2: e = launch_syrk_mode<2>(rows, m, n, col0, s, colmax, partial, st); break;
        default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(syrk_reduce, dim3(s.tiles, 8), dim3(256), 0, st, partial, n, col0, s.tiles, s.splits, colmax,
                       out);
    return hipGetLastError();
}

}  // namespace dsplat
