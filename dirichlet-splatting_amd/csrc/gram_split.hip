// Symmetric rank-k update of pre-split BF16 operands: upper(out) += P^T P, or the rectangle
// out [n, n - col0] += P^T P[:, col0:] (reference: the cuBLAS SSYRK of
// dsplat/kernels/cublas.py:31-44 in varpro.py:91-120, and the streamed cross products of
// replacement.py:178-199).
//
// P [m, n] arrives as two BF16 arrays hi, lo with P ~ hi + lo (hi = RNE(P), lo = RNE(P - hi)),
// or as two FP16 arrays of P with column c scaled by 2^e_c (so the FP16 range holds it; the
// products are unscaled by 2^-(e_i + e_j) in the reduction),
// stored in 32-row column tiles: element (k, c) at ((k / 32) ldt + c) 32 + k % 32 (ldt >= n),
// so the 32 rows of a column are 64 contiguous bytes and the 256 columns of a panel for one
// stage are 16 contiguous KB. The projection kernel writes them in that form (project_td.hip),
// so this kernel does no per-element work besides the products: per 32 rows it stages the two
// 256-column panels of its tile (hi and lo: 64 KB) through LDS and runs hi hi + hi lo + lo hi
// on v_mfma_f32_32x32x16_{bf16,f16} with FP32 accumulation (relative product error about 2^-17
// in BF16, 2^-22 in FP16).
//
// A workgroup (8 waves) owns a 256 x 256 tile and a range of rows (split-K); wave (wi, wj)
// accumulates rows wi * 128 .. + 127 and columns wj * 64 .. + 63 of the tile as 4 x 2 blocks
// of 32 x 32. The next stage's panels load straight into the other LDS buffer while the
// current one is multiplied. In LDS a panel column holds its 32 rows
// as four 16-byte chunks; chunk q of column c sits in slot 4 c + (q ^ ((c >> 2) & 3)), so the
// 16-byte operand reads of 16 consecutive columns cover 16 distinct slots of every 256-byte
// bank window. Partials [split][tile][256][256] are summed in float64 in a fixed order by
// syrk_split_reduce: deterministic, no atomics.
// Lane layouts (gfx950) of 32x32x16: A[i][8 h + e] -> lane i + 32 h, element e; B likewise
// with j; C[8 (r / 4) + 4 (lane / 32) + r % 4][lane % 32] in register r.
#include "dsplat_internal.h"

namespace dsplat {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f32x16 mma_split(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 mma_split(f16x8 a, f16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

constexpr int kGT = 256;        // tile edge
constexpr int kGK = 32;         // rows per stage
constexpr int kGThreads = 512;  // 8 waves
constexpr int kGSlots = kGT * kGK * 2 / 16;  // 16-byte slots per panel part (1024)
constexpr int kGMinSplitRows = 2048;
constexpr int kGMaxSplitRows = 32768;  // rows per float32 partial; the reduction beyond is float64
constexpr int kGCUs = 256;             // one workgroup per CU at a time (128 KB of LDS)

// First output row and column of tile t: the upper triangle (col0 < 0), or the rectangle
// of columns col0 .. n - 1, row-major over tiles.
__device__ __forceinline__ void split_tile_origin(int t, int n, int col0, int* ci, int* cj) {
    const int nb = (n + kGT - 1) / kGT;
    if (col0 < 0) {
        int row = 0;
        while (t >= nb - row) {
            t -= nb - row;
            ++row;
        }
        *ci = row * kGT;
        *cj = (row + t) * kGT;
    } else {
        const int nbj = (n - col0 + kGT - 1) / kGT;
        *ci = (t / nbj) * kGT;
        *cj = col0 + (t % nbj) * kGT;
    }
}

__device__ __forceinline__ int swizzled_slot(int c, int q) { return 4 * c + (q ^ ((c >> 2) & 3)); }

// V: the 8-element operand vector, bf16x8 or f16x8 (16-bit elements either way).
template <typename V>
__global__ __launch_bounds__(kGThreads) void syrk_split_kernel(const uint16_t* __restrict__ hi,
                                                               const uint16_t* __restrict__ lo, long long ldt, int m,
                                                               int n, int col0, int per_split,
                                                               float* __restrict__ partial) {
    extern __shared__ uint4 lds[];  // [2 buffers][4 parts: A hi, A lo, B hi, B lo][kGSlots]
    // the tiles of one row range run on one XCD: the panels they share come through one L2
    const int lin = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int tile = lin % gridDim.x, split = lin / gridDim.x;
    int ci, cj;
    split_tile_origin(tile, n, col0, &ci, &cj);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, h = lane >> 5, l32 = lane & 31;
    const int wi = wave >> 2, wj = wave & 3;
    const int k_begin = split * per_split, k_end = min(m, k_begin + per_split);

    // Staging by the direct global -> LDS loads of gfx950 (global_load_lds_dwordx4: no VGPRs,
    // no wait per load): slot u = tid + 512 it of part p comes from lane tid & 63 of the wave's
    // load it, whose LDS base (M0) is the wave's 64 slots. Slot u holds chunk
    // q = (u & 3) ^ ((c >> 2) & 3) of column c = u >> 2; columns past n read column n - 1 (in
    // bounds; those tile entries are never stored). Every part is loaded, also on a diagonal
    // tile (whose B panel is its A panel, read twice through L2).
    const uint16_t* src[4][2];
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int it = 0; it < 2; ++it) {
            const int u = tid + it * kGThreads, c = u >> 2, q = (u & 3) ^ ((c >> 2) & 3);
            const int col = min((p < 2 ? ci : cj) + c, n - 1);
            src[p][it] = ((p & 1) ? lo : hi) + (size_t)col * 32 + 8 * q;
        }
    const int wave_slot = __builtin_amdgcn_readfirstlane(tid & ~63);
    auto stage = [&](int k0, int b) {
        const size_t off = (size_t)(k0 >> 5) * ldt * 32;
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int it = 0; it < 2; ++it)
                __builtin_amdgcn_global_load_lds(
                    (const __attribute__((address_space(1))) void*)(src[p][it] + off),
                    (__attribute__((address_space(3))) void*)(lds + (b * 4 + p) * kGSlots + it * kGThreads + wave_slot),
                    16, 0, 0);
    };

    f32x16 acc[4][2];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[x][y][r] = 0.0f;

    constexpr int pb = 2;  // B panel parts
    int buf = 0;
    if (k_begin < k_end) stage(k_begin, 0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int k0 = k_begin; k0 < k_end; k0 += kGK) {
        // the next stage's loads go to the other buffer (read by every wave before the last
        // barrier) and stay in flight across the products
        if (k0 + kGK < k_end) stage(k0 + kGK, buf ^ 1);
        const uint4* base = lds + buf * 4 * kGSlots;
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
            const int q = 2 * kk + h;
            V ah[4], al[4], bh[2], bl[2];
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                const int s = swizzled_slot(wi * 128 + x * 32 + l32, q);
                ah[x] = __builtin_bit_cast(V, base[s]);
                al[x] = __builtin_bit_cast(V, base[kGSlots + s]);
            }
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                const int s = swizzled_slot(wj * 64 + y * 32 + l32, q);
                bh[y] = __builtin_bit_cast(V, base[pb * kGSlots + s]);
                bl[y] = __builtin_bit_cast(V, base[(pb + 1) * kGSlots + s]);
            }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y) {
                    f32x16 t = acc[x][y];
                    t = mma_split(ah[x], bh[y], t);
                    t = mma_split(ah[x], bl[y], t);
                    t = mma_split(al[x], bh[y], t);
                    acc[x][y] = t;
                }
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        buf ^= 1;
    }

    float* out = partial + ((size_t)split * gridDim.x + tile) * kGT * kGT;
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int col = wj * 64 + y * 32 + l32;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = wi * 128 + x * 32 + 8 * (r >> 2) + 4 * h + (r & 3);
                out[row * kGT + col] = acc[x][y][r];
            }
        }
}

// out[gi][gj] += sum over splits in float64, fixed order, times 2^-(e_gi + e_gj) when the
// operands carry column exponents: gi <= gj of out [n, n] for the triangle, every entry of
// out [n, n - col0] (column gj - col0) for the rectangle.
__global__ __launch_bounds__(256) void syrk_split_reduce(const float* __restrict__ partial, int n, int col0,
                                                         int tiles, int splits, const float* __restrict__ col_exp,
                                                         float* __restrict__ out) {
    int ci, cj;
    split_tile_origin(blockIdx.x, n, col0, &ci, &cj);
    const int width = col0 < 0 ? n : n - col0, shift = col0 < 0 ? 0 : col0;
    for (int e = blockIdx.y * 256 + threadIdx.x; e < kGT * kGT; e += gridDim.y * 256) {
        const int row = e / kGT, col = e % kGT;
        const int gi = ci + row, gj = cj + col;
        if (gi >= n || gj >= n || (col0 < 0 && gi > gj)) continue;
        double sum = 0.0;
        for (int s = 0; s < splits; ++s) sum += partial[((size_t)s * tiles + blockIdx.x) * kGT * kGT + e];
        if (col_exp) sum = ldexp(sum, -(int)col_exp[gi] - (int)col_exp[gj]);
        out[(size_t)gi * width + (gj - shift)] += (float)sum;
    }
}

SplitShape syrk_split_shape(int m, int n, int col0) {
    SplitShape s;
    const int nb = (n + kGT - 1) / kGT;
    s.tiles = col0 < 0 ? nb * (nb + 1) / 2 : nb * ((n - col0 + kGT - 1) / kGT);
    // One workgroup per CU: as many row ranges as fill whole rounds of the 256 CUs. Float32
    // accumulators sum at most kGMaxSplitRows rows, so long blocks take more rounds; the
    // ranges then fill those rounds.
    const long long fold = ((long long)m + kGMaxSplitRows - 1) / kGMaxSplitRows;
    const long long rounds = (s.tiles * (fold > 1 ? fold : 1) + kGCUs - 1) / kGCUs;
    long long splits = rounds * kGCUs / s.tiles;
    if (splits < fold) splits = fold;
    const int most = (m + kGMinSplitRows - 1) / kGMinSplitRows;
    if (fold <= 1 && splits > most) splits = most;
    if (splits > 65535) splits = 65535;  // grid y
    if (splits < 1) splits = 1;
    s.per_split = (int)((m + splits - 1) / splits);
    s.per_split = (s.per_split + kGK - 1) / kGK * kGK;
    s.splits = (m + s.per_split - 1) / s.per_split;
    if (s.splits < 1) s.splits = 1;
    return s;
}

size_t syrk_split_workspace(int m, int n, int col0) {
    if (m <= 0 || n <= 0 || col0 >= n) return 0;
    const SplitShape s = syrk_split_shape(m, n, col0);
    return sizeof(float) * (size_t)s.splits * s.tiles * kGT * kGT;
}

hipError_t launch_syrk_split(int format, const void* hi, const void* lo, long long ldt, int m, int n, int col0,
                             const float* col_exp, float* out, void* workspace, hipStream_t st) {
    if (m <= 0 || n <= 0 || col0 >= n) return hipSuccess;
    const SplitShape s = syrk_split_shape(m, n, col0);
    float* partial = static_cast<float*>(workspace);
    const size_t smem = 2 * 4 * kGSlots * sizeof(uint4);
    const void* kernel = format == 0 ? (const void*)syrk_split_kernel<bf16x8> : (const void*)syrk_split_kernel<f16x8>;
    hipError_t e = allow_lds(kernel, smem);
    if (e != hipSuccess) return e;
    const uint16_t* h = static_cast<const uint16_t*>(hi);
    const uint16_t* l = static_cast<const uint16_t*>(lo);
    if (format == 0)
        hipLaunchKernelGGL(syrk_split_kernel<bf16x8>, dim3(s.tiles, s.splits), dim3(kGThreads), smem, st, h, l, ldt, m,
                           n, col0, s.per_split, partial);
    else
        hipLaunchKernelGGL(syrk_split_kernel<f16x8>, dim3(s.tiles, s.splits), dim3(kGThreads), smem, st, h, l, ldt, m,
                           n, col0, s.per_split, partial);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(syrk_split_reduce, dim3(s.tiles, 16), dim3(256), 0, st, partial, n, col0, s.tiles, s.splits,
                       col_exp, out);
    return hipGetLastError();
}

}  // namespace dsplat
