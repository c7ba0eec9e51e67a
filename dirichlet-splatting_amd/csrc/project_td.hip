// Row-projected real dictionary blocks in the sample domain (reference: FMCWModel.blocks,
// fmcw.py:243-261: the real bin-major design of surfel.slang:313-378 followed by the
// RowBasis.reduce_major product of fmcw.py:95-98), without writing the design.
//
// With S ADC samples the atom of pair (a, s) at bin k is
//   w D(k - k0) = (w / S) sum_{n<S} e^{-2 pi i n k / N} v^n,   v = e^{2 pi i k0 / N},
// w = sqrt(area) amp rot. A row basis U^H [r, K] applied to the atoms therefore equals the
// sample-domain basis B = U^H F, F[k][n] = e^{-2 pi i n k / N}, applied to the powers of v:
//   (U^H x)_j = (w / S) sum_n B[j][n] v^n.
// sample_basis_kernel forms B in float64 (interleaved real form [2r, 2S], the same row
// layout as the bin-domain basis) once per basis. project_td_kernel then generates
// v^n = Z_{n / 16} W_{n % 16} (one complex product per sample: no reciprocal, no lobe case)
// straight into the B-operand registers of v_mfma_f32_32x32x16_f16 (FP16 hi + lo, three
// products): lane (h, c) of a wave owns surfel column c and the samples 8 t + 4 h .. + 3 of
// every K step t, which is exactly its B fragment. The atoms never touch LDS and the waves
// never wait for each other; only B sits in LDS (FP16 hi/lo), read by every wave.
//
// Work units are (256-column tile, aperture pair). The two lane halves compute the pair
// geometry of the two apertures (one float64 range each) and trade v, v^16 and w through
// v_permlane32_swap. The grid holds as many workgroups as fit on the chip; each takes a
// contiguous range of units.
#include "dsplat_internal.h"

namespace dsplat {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

constexpr int kPtWaves = 8;
constexpr int kPtThreads = 64 * kPtWaves;
constexpr int kPtCols = 32 * kPtWaves;  // surfel columns per unit

// B[row][2 n + p] = sum_k of the interleaved real basis row times the real 2 x 2 block of
// F[k][n] = e^{-2 pi i n k / N}: p = 0 gives Re, p = 1 the second column of the block, so
// row pairs (Re, Im) of U^H map to row pairs of B in the same interleaved layout.
__global__ __launch_bounds__(256) void sample_basis_kernel(const Chirp c, const float* __restrict__ basis,
                                                           float* __restrict__ out) {
    extern __shared__ double2 T[];  // (cos, sin) of 2 pi m / N
    for (int m = threadIdx.x; m < c.N; m += 256) {
        double s, co;
        sincospi(2.0 * (double)m / (double)c.N, &s, &co);
        T[m] = make_double2(co, s);
    }
    __syncthreads();
    const int col = blockIdx.x * 256 + threadIdx.x;
    if (col >= 2 * c.S) return;
    const int n = col >> 1, p = col & 1;
    const float* b = basis + (size_t)blockIdx.y * 2 * c.K;
    double acc = 0.0;
    int m = 0;  // n k mod N
    for (int k = 0; k < c.K; ++k) {
        const double2 e = T[m];
        const double br = b[2 * k], bi = b[2 * k + 1];
        acc += p ? fma(br, e.y, bi * e.x) : fma(br, e.x, -bi * e.y);
        m += n;
        m -= m >= c.N ? c.N : 0;
    }
    out[(size_t)blockIdx.y * 2 * c.S + col] = (float)acc;
}

__device__ __forceinline__ f32x16 mfma_f16(f16x8 a, f16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

// Where a projected block goes. Row-major float32 out [rows * A, S] (row j * A + a), or
// (split) BF16 or FP16 hi / lo with hi + lo ~ value (FP16: value times 2^col_exp[s]) in
// 32-row column tiles, element (k, s) at
// ((k / 32) ld + s) 32 + k % 32, where projected row j of aperture a is k = a * rows_pad + j,
// rows_pad = rows rounded up to 32 (rows j >= rows hold zeros): the operand form of
// syrk_split_kernel.
struct ProjectOut {
    float* out;
    unsigned* colmax;
    void* hi;              // split: BF16 or FP16 [m / 32][ld][32]
    void* lo;
    long long ld;
    int rows_pad;
    const float* col_exp;  // split FP16: column s is scaled by 2^col_exp[s] before the split
};

// Output forms of project_td_kernel.
constexpr int kOutFloat = 0;  // row-major float32
constexpr int kOutBf16 = 1;   // split BF16 hi / lo
constexpr int kOutFp16 = 2;   // split FP16 hi / lo of the column-scaled value

template <typename T>
__device__ __forceinline__ void split_store(float x[4], void* hi, void* lo, size_t at) {
    T vh[4], vl[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        vh[e] = (T)x[e];
        vl[e] = (T)(x[e] - (float)vh[e]);
    }
    *reinterpret_cast<uint2*>(static_cast<T*>(hi) + at) = __builtin_bit_cast(uint2, vh);
    *reinterpret_cast<uint2*>(static_cast<T*>(lo) + at) = __builtin_bit_cast(uint2, vl);
}

// One aperture of one unit: v, z = v^16 and the weight w / S of this lane's pair (times the
// column's power-of-two scale in the FP16 split form). Returns the largest |value| this lane
// stored (0 for a padding column).
template <int J, int NB, int OUT>
__device__ __forceinline__ float project_aperture(const _Float16* Bh, const _Float16* Bl, int bs, int h, int l32,
                                                 float2 v, float2 z, float2 w, int row0, int rows, int a, int A,
                                                 int s, int S, const ProjectOut& o) {
    // W of this lane's residues i = 4 h + e (even K steps) and 8 + 4 h + e (odd steps)
    const float2 v2 = csq(v), v4 = csq(v2), v8 = csq(v4);
    float2 W[2][4];
    W[0][0] = h ? v4 : make_float2(1.f, 0.f);
    W[0][1] = cmul(W[0][0], v);
    W[0][2] = cmul(W[0][0], v2);
    W[0][3] = cmul(W[0][1], v2);
#pragma unroll
    for (int e = 0; e < 4; ++e) W[1][e] = cmul(W[0][e], v8);
    // Z_j = z^j by a binary product tree
    float2 Z[J];
    Z[0] = make_float2(1.f, 0.f);
    Z[1] = z;
#pragma unroll
    for (int j = 2; j < J; ++j) Z[j] = (j & 1) ? cmul(Z[j - 1], z) : csq(Z[j >> 1]);

    f32x16 acc[NB];
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[x][r] = 0.0f;
    // basis fragments of step t + 1 are read while step t runs (two register sets); the
    // scheduling barrier at the end of a step keeps later reads from piling up
    f16x8 uh[2][NB], ul[2][NB];
    auto fetch = [&](int t, int buf) {
#pragma unroll
        for (int x = 0; x < NB; ++x) {
            const int off = (x * 32 + l32) * bs + t * 16 + 8 * h;
            uh[buf][x] = *reinterpret_cast<const f16x8*>(Bh + off);
            ul[buf][x] = *reinterpret_cast<const f16x8*>(Bl + off);
        }
    };
    fetch(0, 0);
#pragma unroll
    for (int t = 0; t < 2 * J; ++t) {
        const int b = t & 1;
        if (t + 1 < 2 * J) fetch(t + 1, b ^ 1);
        // samples 8 t + 4 h + e = 16 (t / 2) + i: v^n = Z_{t / 2} W_i, as (re, im) pairs
        f16x8 xh, xl;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float2 p = cmul(Z[t >> 1], W[t & 1][e]);
            f16x2 ph, pl;
            split_f16_pair<true>(p.x, p.y, ph, pl);
            xh[2 * e] = ph.x, xh[2 * e + 1] = ph.y;
            xl[2 * e] = pl.x, xl[2 * e + 1] = pl.y;
        }
#pragma unroll
        for (int x = 0; x < NB; ++x) acc[x] = mfma_f16(uh[b][x], xh, acc[x]);
#pragma unroll
        for (int x = 0; x < NB; ++x) acc[x] = mfma_f16(uh[b][x], xl, acc[x]);
#pragma unroll
        for (int x = 0; x < NB; ++x) acc[x] = mfma_f16(ul[b][x], xh, acc[x]);
        __builtin_amdgcn_sched_barrier(0);
    }
    // (Re, Im) row pairs times w; C[row][col] -> acc[r] with row = 8 (r / 4) + 4 h + r % 4, so
    // rows 2j and 2j + 1 are acc[r], acc[r + 1] for even r
    if (s >= S) return 0.0f;
    if constexpr (OUT != kOutFloat) {
        // rows 4 g .. 4 g + 3 of a 32-row block are (re, im) of two projected rows: 8 bytes of
        // hi and of lo per group; the two lane halves fill 16 contiguous bytes, the four groups
        // the column's 64 bytes of the tile
        const size_t tile0 = ((size_t)a * o.rows_pad + row0) / 32;
#pragma unroll
        for (int x = 0; x < NB; ++x) {
            if (row0 + x * 32 >= o.rows_pad) break;  // uniform
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                float val[4];
#pragma unroll
                for (int e = 0; e < 4; e += 2) {
                    const float hr = acc[x][4 * g + e], hm = acc[x][4 * g + e + 1];
                    val[e] = w.x * hr - w.y * hm;
                    val[e + 1] = w.x * hm + w.y * hr;
                }
                const size_t at = ((tile0 + x) * o.ld + s) * 32 + 8 * g + 4 * h;
                if constexpr (OUT == kOutBf16)
                    split_store<__bf16>(val, o.hi, o.lo, at);
                else
                    split_store<_Float16>(val, o.hi, o.lo, at);
            }
        }
        return 0.0f;
    }
    float* __restrict__ out = o.out;
    // The row stride passes through an empty asm so that the 64-bit row offsets are formed
    // here, per aperture, in scalar registers, instead of being hoisted out of every loop
    // into dozens of live registers.
    unsigned long long stride = (unsigned long long)A * S;
    asm volatile("" : "+s"(stride));
    float* p = out + ((size_t)(row0 + 4 * h) * A + a) * S + s;
    const bool full = row0 + 32 * NB <= rows;  // uniform: no per-row test on whole passes
    float top = 0.0f;
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int r = 0; r < 16; r += 2) {
            const int dr = x * 32 + 8 * (r >> 2) + (r & 3);  // row - row0 - 4 h, even
            if (full || row0 + 4 * h + dr < rows) {           // rows is even: row + 1 < rows too
                const float hr = acc[x][r], hm = acc[x][r + 1];
                const float re = w.x * hr - w.y * hm, im = w.x * hm + w.y * hr;
                p[dr * stride] = re;
                p[(dr + 1) * stride] = im;
                top = fmaxf(top, fmaxf(fabsf(re), fabsf(im)));
            }
        }
    return top;
}

// colmax[s] = max(colmax[s], top of both lane halves) as float bits (non-negative floats
// order like unsigned ints).
__device__ __forceinline__ void flush_column_max(unsigned* colmax, float top, int h, int s, int S) {
    float other = top;
    swap_halves(top, other);  // top: half 0's value in every lane, other: half 1's
    if (colmax && h == 0 && s < S) atomicMax(colmax + s, __float_as_uint(fmaxf(top, other)));
}

template <int J, bool FRESNEL, int ROWS, int OUT>
__global__ __launch_bounds__(kPtThreads) void project_td_kernel(const Chirp c, const float* __restrict__ pos, int A,
                                                                const float* __restrict__ ctr,
                                                                const float* __restrict__ nrm,
                                                                const float* __restrict__ areas, int S,
                                                                const float* __restrict__ basis, int rows,
                                                                const ProjectOut o) {
    unsigned* __restrict__ colmax = o.colmax;
    constexpr int KS = 32 * J;  // real values per basis row: (re, im) of 16 J samples
    constexpr int BS = KS + 8;  // LDS row stride in halves
    constexpr int NB = ROWS / 32;
    extern __shared__ __attribute__((aligned(16))) float2 E[];  // [N] twiddles
    _Float16* Bh = reinterpret_cast<_Float16*>(E + c.N);       // [ROWS][BS]
    _Float16* Bl = Bh + ROWS * BS;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, h = lane >> 5, l32 = lane & 31;
    const int row0 = blockIdx.y * ROWS;

    fill_twiddles(E, c.N, tid, kPtThreads);
    for (int i = tid; i < ROWS * KS / 2; i += kPtThreads) {
        const int row = i / (KS / 2), kk = 2 * (i - row * (KS / 2));
        float2 b = make_float2(0.f, 0.f);
        if (row0 + row < rows) b = *reinterpret_cast<const float2*>(basis + (size_t)(row0 + row) * KS + kk);
        f16x2 hi, lo;
        split_f16_pair<false>(b.x, b.y, hi, lo);
        *reinterpret_cast<f16x2*>(Bh + row * BS + kk) = hi;
        *reinterpret_cast<f16x2*>(Bl + row * BS + kk) = lo;
    }
    __syncthreads();

    // Units in contiguous ranges per workgroup, ordered tile-major: a lane's surfel (vector
    // loads) changes only with the tile, and the apertures come through scalar loads. Vector
    // loads inside the unit loop would make each unit wait for all stores of the previous one
    // (one in-order counter covers both).
    const int tiles = (S + kPtCols - 1) / kPtCols, pairs = (A + 1) / 2;
    const int units = tiles * pairs;
    const int per = (units + gridDim.x - 1) / gridDim.x;
    const int u_end = min(units, (int)blockIdx.x * per + per);
    int tile = -1, s = 0;
    float cx = 0.f, cy = 0.f, cz = 0.f, nx = 0.f, ny = 0.f, nz = 0.f, root_area = 0.f;
    float column_scale = 1.0f;  // FP16 split: 2^col_exp of this lane's column
    float top = 0.0f;  // largest |value| of this lane's column in the current tile
    for (int u = blockIdx.x * per; u < u_end; ++u) {
        if (u / pairs != tile) {
            if (tile >= 0) flush_column_max(colmax, top, h, s, S);
            top = 0.0f;
            tile = u / pairs;
            s = tile * kPtCols + wave * 32 + l32;
            const int sc = min(s, S - 1);
            cx = ctr[3 * sc], cy = ctr[3 * sc + 1], cz = ctr[3 * sc + 2];
            nx = nrm[3 * sc], ny = nrm[3 * sc + 1], nz = nrm[3 * sc + 2];
            root_area = sqrtf(areas[sc]);
            if constexpr (OUT == kOutFp16) column_scale = ldexpf(1.0f, (int)o.col_exp[sc]);
            __builtin_amdgcn_s_waitcnt(0);  // here, not at the first use that every unit shares
        }
        const int a0 = 2 * (u - tile * pairs), a1 = min(a0 + 1, A - 1);
        // lane half h: aperture a0 + h (both positions by uniform loads, then a select)
        const float p0x = pos[3 * a0], p0y = pos[3 * a0 + 1], p0z = pos[3 * a0 + 2];
        const float p1x = pos[3 * a1], p1y = pos[3 * a1 + 1], p1z = pos[3 * a1 + 2];
        const float px = h ? p1x : p0x, py = h ? p1y : p0y, pz = h ? p1z : p0z;
        const Pair q = pair_geometry_t<FRESNEL>(c, px, py, pz, cx, cy, cz, nx, ny, nz);
        const float2 v = rotation(c, E, q.lobe, q.delta, 1);
        const float2 z = rotation(c, E, q.lobe, q.delta, 16);
        const float2 w = (q.amp * root_area * c.inv_S * column_scale) * q.rot;
        // every lane gets both apertures' values: f0 from half 0 (a0), f1 from half 1 (a0 + 1)
        float f0[6] = {v.x, v.y, z.x, z.y, w.x, w.y}, f1[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            f1[i] = f0[i];
            swap_halves(f0[i], f1[i]);
        }
#pragma unroll 1
        for (int e = 0; e < 2 && a0 + e < A; ++e)
            top = fmaxf(top, project_aperture<J, NB, OUT>(Bh, Bl, BS, h, l32,
                                                            make_float2(e ? f1[0] : f0[0], e ? f1[1] : f0[1]),
                                                            make_float2(e ? f1[2] : f0[2], e ? f1[3] : f0[3]),
                                                            make_float2(e ? f1[4] : f0[4], e ? f1[5] : f0[5]), row0,
                                                            rows, a0 + e, A, s, S, o));
    }
    if (tile >= 0) flush_column_max(colmax, top, h, s, S);
}

static size_t project_td_lds(const Chirp& c, int rows_tile) {
    return sizeof(float2) * c.N + 2 * sizeof(_Float16) * (size_t)rows_tile * (2 * c.S + 8);
}

// The row tile: 128 basis rows per pass where they fit and are needed, else 64.
static int project_td_tile(const Chirp& c, int rows) {
    if (rows > 64 && project_td_lds(c, 128) <= 160 * 1024) return 128;
    return project_td_lds(c, 64) <= 160 * 1024 ? 64 : 0;
}

bool project_td_supported(const Chirp& c, int rows) {
    return use_td(c) && rows >= 2 && rows % 2 == 0 && project_td_tile(c, rows) > 0;
}

template <int J, bool F, int ROWS, int OUT>
static hipError_t launch_ptd(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                             const float* areas, int S, const float* basis, int rows, const ProjectOut& o,
                             hipStream_t st) {
    const void* kernel = (const void*)project_td_kernel<J, F, ROWS, OUT>;
    const size_t smem = project_td_lds(c, ROWS);
    hipError_t e = allow_lds(kernel, smem);
    if (e != hipSuccess) return e;
    int dev = 0, cus = 0, occ = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, kPtThreads, smem)) != hipSuccess) return e;
    const int passes = (rows + ROWS - 1) / ROWS;
    const long long units = (long long)((S + kPtCols - 1) / kPtCols) * ((A + 1) / 2);
    long long blocks = (long long)cus * (occ > 0 ? occ : 1) / passes;
    if (blocks < 1) blocks = 1;
    if (blocks > units) blocks = units;
    hipLaunchKernelGGL((project_td_kernel<J, F, ROWS, OUT>), dim3((unsigned)blocks, passes), dim3(kPtThreads),
                       smem, st, c, pos, A, ctr, nrm, areas, S, basis, rows, o);
    return hipGetLastError();
}

template <int J, bool F, int OUT>
static hipError_t launch_ptd_j(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                               const float* areas, int S, const float* basis, int rows, const ProjectOut& o,
                               hipStream_t st) {
    if constexpr (J == 8) {
        if (project_td_tile(c, rows) == 128)
            return launch_ptd<J, F, 128, OUT>(c, pos, A, ctr, nrm, areas, S, basis, rows, o, st);
    }
    return launch_ptd<J, F, 64, OUT>(c, pos, A, ctr, nrm, areas, S, basis, rows, o, st);
}

template <int OUT>
static hipError_t launch_ptd_any(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                                 const float* areas, int S, const float* basis, int rows, const ProjectOut& o,
                                 hipStream_t st) {
    const bool fresnel = c.fresnel_n > 1.0f;
    if (c.S == 128)
        return fresnel ? launch_ptd_j<8, true, OUT>(c, pos, A, ctr, nrm, areas, S, basis, rows, o, st)
                       : launch_ptd_j<8, false, OUT>(c, pos, A, ctr, nrm, areas, S, basis, rows, o, st);
    return fresnel ? launch_ptd_j<16, true, OUT>(c, pos, A, ctr, nrm, areas, S, basis, rows, o, st)
                   : launch_ptd_j<16, false, OUT>(c, pos, A, ctr, nrm, areas, S, basis, rows, o, st);
}

hipError_t launch_sample_basis(const Chirp& c, const float* basis, int rows, float* out, hipStream_t st) {
    hipLaunchKernelGGL(sample_basis_kernel, dim3((2 * c.S + 255) / 256, rows), dim3(256), sizeof(double2) * c.N, st,
                       c, basis, out);
    return hipGetLastError();
}

hipError_t launch_project_td(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                             const float* areas, int S, const float* basis, int rows, float* out, unsigned* colmax,
                             hipStream_t st) {
    ProjectOut o{out, colmax, nullptr, nullptr, 0, 0, nullptr};
    return launch_ptd_any<kOutFloat>(c, pos, A, ctr, nrm, areas, S, basis, rows, o, st);
}

int project_rows_pad(int rows) { return (rows + 31) / 32 * 32; }

hipError_t launch_project_split(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                                const float* areas, int S, const float* basis, int rows, int format,
                                const float* col_exp, void* hi, void* lo, long long ld, hipStream_t st) {
    ProjectOut o{nullptr, nullptr, hi, lo, ld, project_rows_pad(rows), col_exp};
    if (format == 0) return launch_ptd_any<kOutBf16>(c, pos, A, ctr, nrm, areas, S, basis, rows, o, st);
    return launch_ptd_any<kOutFp16>(c, pos, A, ctr, nrm, areas, S, basis, rows, o, st);
}

}  // namespace dsplat
