// Fused Gram G = A^H A of the surfel dictionary on MFMA, atoms generated on chip
// (reference: the streamed normal equations of dsplat/varpro.py:91-120 with the design of
// kernels/surfel.slang:313-378 and the cuBLAS SSYRK of kernels/cublas.py).
//
// A[(a, k), s] = sqrt(area_s) amp rot D(k - k0) for every aperture a and bin k (rows) and
// surfel s (columns). A workgroup owns a 128 x 128 complex tile (I <= J) of G and a range
// of apertures (split-K). Per aperture each thread owns one column: its float64 pair
// prologue stays in registers while it generates 8 bins per K-step into LDS. Each wave
// accumulates a 64 x 64 complex sub-tile as 2 x 2 MFMA blocks of 32 x 32:
//   Re G += X^T Y,  Im G += X^T Y',   with K-order k = part * 8 + bin, Y' = (im, -re).
// MODE 0: FP32 operands, v_mfma_f32_32x32x2f32.
// MODE 1: BF16 hi + lo split, 3 v_mfma_f32_32x32x16_bf16 per product (hi hi + hi lo + lo hi).
// MODE 2: FP16 hi + lo split, same with _f16, columns scaled by 2^e_s into FP16 range.
// Partials [split][S][S] complex (upper tiles only) are summed in float64 in a fixed order,
// unscaled, and mirrored to the full Hermitian matrix by gram_reduce.
#include "dsplat_atoms.h"
#include "dsplat_internal.h"

namespace dsplat {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int kGramTile = 128;  // complex columns per tile side
constexpr int kGramStep = 8;    // bins per K-step (16 real K values)
constexpr int kFoldApertures = 128;  // apertures per float32 partial (float64 fold beyond)

template <int MODE>
struct GramTypes;
template <>
struct GramTypes<1> {
    typedef __bf16 scalar;
    typedef bf16x8 vec;
};
template <>
struct GramTypes<2> {
    typedef _Float16 scalar;
    typedef f16x8 vec;
};

__device__ __forceinline__ f32x16 mfma16(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 mfma16(f16x8 a, f16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

// Negate every element of an 8-wide 16-bit vector (sign-bit flip), for the (im, -re) operand.
template <typename V>
__device__ __forceinline__ V negate_if(V v, bool neg) {
    u32x4 bits = __builtin_bit_cast(u32x4, v);
    const unsigned mask = neg ? 0x80008000u : 0u;
    bits ^= mask;
    return __builtin_bit_cast(V, bits);
}

template <int P, int MODE>
__global__ __launch_bounds__(256) void gram_kernel(const Chirp c, const float* __restrict__ pos, int A,
                                                   const float* __restrict__ ctr, const float* __restrict__ nrm,
                                                   const float* __restrict__ areas,
                                                   const float* __restrict__ col_exp, int S, int real_only,
                                                   int per_split, float* __restrict__ partial) {
    extern __shared__ float smem[];
    float* T = smem;  // [2N]
    const int nb = (S + kGramTile - 1) / kGramTile;
    int bi = 0, t = blockIdx.x;
    while (t >= nb - bi) {
        t -= nb - bi;
        ++bi;
    }
    const int bj = bi + t;
    const bool diagonal = bi == bj;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wi = wave >> 1, wj = wave & 1;
    const int split = blockIdx.y;
    const int a_begin = split * per_split, a_end = min(A, a_begin + per_split);
    // Thread -> generated column: 0..127 the I block, 128..255 the J block (idle on the diagonal).
    const bool gen = !(diagonal && tid >= kGramTile);
    const int slot = tid & (kGramTile - 1);
    const int s = (tid < kGramTile ? bi : bj) * kGramTile + slot;
    const float e = s < S && col_exp ? col_exp[s] : 0.0f;
    const float scale = ldexpf(1.0f, (int)e);

    for (int i = tid; i < 2 * c.N; i += 256) T[i] = cot_entry(i, c.N, 0.0f);
    __syncthreads();

    // LDS operand buffers [2 buffers][2 blocks I/J][128 columns][16 K values] (x2 for hi/lo).
    constexpr int kVals = 2 * kGramStep;
    f32x16 acc_re[2][2], acc_im[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                acc_re[x][y][r] = 0.0f;
                acc_im[x][y][r] = 0.0f;
            }

    if constexpr (MODE == 0) {
        float* buf = T + 2 * c.N;  // [2][2][128][16] floats
        auto at = [&](int b, int blk, int colx, int k) -> float& {
            return buf[(((b * 2 + blk) * kGramTile + colx) * kVals) + k];
        };
        int step = 0;
        for (int a = a_begin; a < a_end; ++a) {
            const Column<P> col = column_state<P>(c, pos, a, ctr, nrm, areas, scale, s, S);
            for (int k0 = 0; k0 < c.K; k0 += kGramStep, ++step) {
                const int b = step & 1;
                if (gen) {
                    const int blk = tid < kGramTile ? 0 : 1;
#pragma unroll
                    for (int j = 0; j < kGramStep; ++j) {
                        const int k = k0 + j;
                        const float2 v = column_atom<P>(c, T, col, k, k % P);
                        at(b, blk, slot, j) = v.x;
                        at(b, blk, slot, kGramStep + j) = v.y;
                    }
                }
                __syncthreads();
                const int jblk = diagonal ? 0 : 1;
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int y = 0; y < 2; ++y) {
                        const int ci = wi * 64 + x * 32 + (lane & 31);
                        const int cj = wj * 64 + y * 32 + (lane & 31);
#pragma unroll
                        for (int kk = 0; kk < kVals; kk += 2) {
                            const int k = kk + (lane >> 5);
                            const float xa = at(b, 0, ci, k);
                            const float yb = at(b, jblk, cj, k);
                            acc_re[x][y] = __builtin_amdgcn_mfma_f32_32x32x2f32(xa, yb, acc_re[x][y], 0, 0, 0);
                            if (!real_only) {
                                // Y'[k] = im (k < 8) or -re (k >= 8) of the same bin; kk picks the half
                                const float yi = at(b, jblk, cj, (k + kGramStep) & (kVals - 1));
                                acc_im[x][y] = __builtin_amdgcn_mfma_f32_32x32x2f32(
                                    xa, kk < kGramStep ? yi : -yi, acc_im[x][y], 0, 0, 0);
                            }
                        }
                    }
            }
        }
    } else {
        typedef typename GramTypes<MODE>::scalar h16;
        typedef typename GramTypes<MODE>::vec v8;
        h16* buf = reinterpret_cast<h16*>(T + 2 * c.N);  // [2][2 hi/lo][2 blocks][128][16]
        auto idx = [&](int b, int part, int blk, int colx, int k) {
            return ((((b * 2 + part) * 2 + blk) * kGramTile + colx) * kVals) + k;
        };
        int step = 0;
        for (int a = a_begin; a < a_end; ++a) {
            const Column<P> col = column_state<P>(c, pos, a, ctr, nrm, areas, scale, s, S);
            for (int k0 = 0; k0 < c.K; k0 += kGramStep, ++step) {
                const int b = step & 1;
                if (gen) {
                    const int blk = tid < kGramTile ? 0 : 1;
                    v8 hre, lre, him, lim;
#pragma unroll
                    for (int j = 0; j < kGramStep; ++j) {
                        const int k = k0 + j;
                        const float2 v = column_atom<P>(c, T, col, k, k % P);
                        const h16 hr = (h16)v.x, hi = (h16)v.y;
                        hre[j] = hr;
                        him[j] = hi;
                        lre[j] = (h16)(v.x - (float)hr);
                        lim[j] = (h16)(v.y - (float)hi);
                    }
                    *reinterpret_cast<v8*>(&buf[idx(b, 0, blk, slot, 0)]) = hre;
                    *reinterpret_cast<v8*>(&buf[idx(b, 0, blk, slot, kGramStep)]) = him;
                    *reinterpret_cast<v8*>(&buf[idx(b, 1, blk, slot, 0)]) = lre;
                    *reinterpret_cast<v8*>(&buf[idx(b, 1, blk, slot, kGramStep)]) = lim;
                }
                __syncthreads();
                const int jblk = diagonal ? 0 : 1;
                const int half = lane >> 5;
#pragma unroll
                for (int x = 0; x < 2; ++x) {
                    const int ci = wi * 64 + x * 32 + (lane & 31);
                    const v8 xh = *reinterpret_cast<const v8*>(&buf[idx(b, 0, 0, ci, 8 * half)]);
                    const v8 xl = *reinterpret_cast<const v8*>(&buf[idx(b, 1, 0, ci, 8 * half)]);
#pragma unroll
                    for (int y = 0; y < 2; ++y) {
                        const int cj = wj * 64 + y * 32 + (lane & 31);
                        const v8 yh = *reinterpret_cast<const v8*>(&buf[idx(b, 0, jblk, cj, 8 * half)]);
                        const v8 yl = *reinterpret_cast<const v8*>(&buf[idx(b, 1, jblk, cj, 8 * half)]);
                        f32x16 re = acc_re[x][y];
                        re = mfma16(xh, yh, re);
                        re = mfma16(xh, yl, re);
                        re = mfma16(xl, yh, re);
                        acc_re[x][y] = re;
                        if (!real_only) {
                            // Y' for lane half h is the other half of the same column, negated on h = 1
                            const v8 ph = negate_if(*reinterpret_cast<const v8*>(&buf[idx(b, 0, jblk, cj, 8 * (1 - half))]), half);
                            const v8 pl = negate_if(*reinterpret_cast<const v8*>(&buf[idx(b, 1, jblk, cj, 8 * (1 - half))]), half);
                            f32x16 im = acc_im[x][y];
                            im = mfma16(xh, ph, im);
                            im = mfma16(xh, pl, im);
                            im = mfma16(xl, ph, im);
                            acc_im[x][y] = im;
                        }
                    }
                }
            }
        }
    }
    // partial[split][gi][gj] (complex) for the computed tile; C[row][col] -> acc[r],
    // row = 8 (r / 4) + 4 (lane / 32) + r % 4, col = lane % 32.
    float2* out = reinterpret_cast<float2*>(partial) + (size_t)split * S * S;
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int gj = bj * kGramTile + wj * 64 + y * 32 + (lane & 31);
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int gi = bi * kGramTile + wi * 64 + x * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (gi < S && gj < S) out[(size_t)gi * S + gj] = make_float2(acc_re[x][y][r], acc_im[x][y][r]);
            }
        }
}

// G[i][j] = sum over splits of partial[.][i][j] (i <= j in tile order) in float64, unscaled
// by 2^-(e_i + e_j); the other triangle is the conjugate. real_only writes Re into a
// float [S, S] matrix, else complex [S, S].
__global__ __launch_bounds__(256) void gram_reduce(const float* __restrict__ partial, int S, int splits,
                                                   const float* __restrict__ col_exp, int real_only,
                                                   float inv_rows, float* __restrict__ out) {
    const size_t idx = blockIdx.x * (size_t)256 + threadIdx.x;
    if (idx >= (size_t)S * S) return;
    const int i = (int)(idx / S), j = (int)(idx % S);
    const int ti = i / kGramTile, tj = j / kGramTile;
    // tiles with ti <= tj hold (i, j) directly; below the diagonal of tiles use (j, i)
    const bool direct = ti < tj || (ti == tj && i <= j);
    const int r = direct ? i : j, q = direct ? j : i;
    const float2* p = reinterpret_cast<const float2*>(partial);
    double re = 0.0, im = 0.0;
    for (int sp = 0; sp < splits; ++sp) {
        const float2 v = p[((size_t)sp * S + r) * S + q];
        re += v.x;
        im += v.y;
    }
    const double unscale = col_exp ? ldexp(1.0, -(int)col_exp[i] - (int)col_exp[j]) : 1.0;
    re *= unscale * inv_rows;
    im *= unscale * inv_rows;
    if (!direct) im = -im;
    if (i == j) im = 0.0;  // the diagonal of A^H A is real
    if (real_only) out[idx] = (float)re;
    else reinterpret_cast<float2*>(out)[idx] = make_float2((float)re, (float)im);
}

GramShape gram_shape(int A, int S) {
    // Split-K over apertures: at most kFoldApertures apertures (128 x K rows) accumulate in
    // one float32 partial before the float64 reduction, and enough ranges to give ~4
    // workgroups per CU.
    GramShape g;
    const int nb = (S + kGramTile - 1) / kGramTile;
    g.tiles = nb * (nb + 1) / 2;
    int splits = (1024 + g.tiles - 1) / g.tiles;
    const int fold = (A + kFoldApertures - 1) / kFoldApertures;
    if (splits < fold) splits = fold;
    if (splits > A) splits = A;
    if (splits < 1) splits = 1;
    g.per_split = (A + splits - 1) / splits;
    g.splits = (A + g.per_split - 1) / g.per_split;
    return g;
}

size_t gram_smem(const Chirp& c, int mode) {
    const size_t ops = mode == 0 ? sizeof(float) * 2 * 2 * kGramTile * 2 * kGramStep
                                 : sizeof(short) * 2 * 2 * 2 * kGramTile * 2 * kGramStep;
    return sizeof(float) * 2 * c.N + ops;
}

template <int P, int MODE>
static hipError_t launch_gram_pm(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                                 const float* areas, const float* col_exp, int S, int real_only, float* partial,
                                 hipStream_t st) {
    const GramShape g = gram_shape(A, S);
    const size_t smem = gram_smem(c, MODE);
    hipError_t e = allow_lds((const void*)gram_kernel<P, MODE>, smem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((gram_kernel<P, MODE>), dim3(g.tiles, g.splits), dim3(256), smem, st, c, pos, A, ctr, nrm,
                       areas, col_exp, S, real_only, g.per_split, partial);
    return hipGetLastError();
}

template <int P>
static hipError_t launch_gram_p(const Chirp& c, int mode, const float* pos, int A, const float* ctr,
                                const float* nrm, const float* areas, const float* col_exp, int S, int real_only,
                                float* partial, hipStream_t st) {
    switch (mode) {
        case 0: return launch_gram_pm<P, 0>(c, pos, A, ctr, nrm, areas, col_exp, S, real_only, partial, st);
        case 1: return launch_gram_pm<P, 1>(c, pos, A, ctr, nrm, areas, col_exp, S, real_only, partial, st);
        case 2: return launch_gram_pm<P, 2>(c, pos, A, ctr, nrm, areas, col_exp, S, real_only, partial, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_gram(const Chirp& c, int mode, const float* pos, int A, const float* ctr, const float* nrm,
                       const float* areas, const float* col_exp, int S, int real_only, float inv_rows,
                       float* partial, float* out, hipStream_t st) {
    hipError_t e;
    switch (c.P) {
        case 1: e = launch_gram_p<1>(c, mode, pos, A, ctr, nrm, areas, col_exp, S, real_only, partial, st); break;
        case 2: e = launch_gram_p<2>(c, mode, pos, A, ctr, nrm, areas, col_exp, S, real_only, partial, st); break;
        case 4: e = launch_gram_p<4>(c, mode, pos, A, ctr, nrm, areas, col_exp, S, real_only, partial, st); break;
        case 8: e = launch_gram_p<8>(c, mode, pos, A, ctr, nrm, areas, col_exp, S, real_only, partial, st); break;
        default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    const GramShape g = gram_shape(A, S);
    const size_t total = (size_t)S * S;
    hipLaunchKernelGGL(gram_reduce, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, partial, S, g.splits,
                       col_exp, real_only, inv_rows, out);
    return hipGetLastError();
}

}  // namespace dsplat
