// Row-projected real dictionary blocks in one pass (reference: FMCWModel.blocks,
// fmcw.py:243-261: the real bin-major design of surfel.slang:313-378 followed by the
// RowBasis.reduce_major product of fmcw.py:95-98), without writing the design.
//
//   out[(j A + a) S + s] = sum over (k, part) of basis[j][2 k + part] x_part(a, k, s),
//
// x = (Re, Im) of atom (a, k, s) = w_{a,s} D(k - k0_{a,s}), w = sqrt(area) amp rot, and basis
// the interleaved real form [2r, 2K] of U^H (rows 2j, 2j+1 give Re and Im of (U^H x)_j). As w
// does not depend on k, the unit-weight atoms D are projected on MFMA and each (Re, Im) row
// pair is multiplied by w afterwards, so every MFMA operand is O(1) and FP16 hi + lo keeps
// about 22 bits.
//
// A workgroup owns 128 surfel columns and a range of apertures. The basis sits in LDS as
// FP16 hi/lo, k-contiguous per row; per aperture, every K-step of 8 bins (16 real values) is
// generated into a double-buffered LDS tile (two threads per column, 4 bins each) and
// contracted with v_mfma_f32_32x32x16_f16 three times (hi hi + hi lo + lo hi). Wave w owns
// output columns 32w..32w+31 and every basis row (ROWS / 32 blocks).
#include "dsplat_atoms.h"
#include "dsplat_internal.h"

namespace dsplat {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

constexpr int kProjCols = 128;
constexpr int kProjStep = 16;  // real K values per MFMA step (8 bins)

__device__ __forceinline__ f32x16 mfma_f16(f16x8 a, f16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}

template <int P, int ROWS>
__global__ __launch_bounds__(256) void project_kernel(const Chirp c, const float* __restrict__ pos, int A,
                                                      const float* __restrict__ ctr, const float* __restrict__ nrm,
                                                      const float* __restrict__ areas, int S,
                                                      const float* __restrict__ basis, int rows, int per_group,
                                                      float* __restrict__ out) {
    extern __shared__ __attribute__((aligned(16))) float smem[];
    const int steps = (c.K + 7) / 8;          // 8 bins (16 real values) per MFMA step
    const int KK = steps * kProjStep;         // real K padded to whole steps
    const int basis_stride = KK + 8;          // 16-bit values per basis row in LDS
    const int row0 = blockIdx.z * ROWS;       // this pass's first basis row
    float* T = smem;                         // [2N]
    float2* weight = reinterpret_cast<float2*>(T + 2 * c.N);  // [128] w of the current aperture
    _Float16* Bh = reinterpret_cast<_Float16*>(weight + kProjCols);  // basis hi [ROWS][basis_stride]
    _Float16* Bl = Bh + ROWS * basis_stride;                         // basis lo
    _Float16* X = Bl + ROWS * basis_stride;  // atoms [2 buf][2 hi/lo][128 cols][kProjStep + 8]
    constexpr int xs = kProjStep + 8;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, h = lane >> 5, l32 = lane & 31;
    const int s0 = blockIdx.x * kProjCols;
    const int a_begin = blockIdx.y * per_group, a_end = min(A, a_begin + per_group);

    for (int i = tid; i < 2 * c.N; i += 256) T[i] = cot_entry(i, c.N, 0.0f);
    for (int i = tid; i < ROWS * KK; i += 256) {
        const int row = i / KK, kk = i - row * KK;
        const float v = row0 + row < rows && kk < 2 * c.K ? basis[(size_t)(row0 + row) * 2 * c.K + kk] : 0.0f;
        const _Float16 hi = (_Float16)v;
        Bh[row * basis_stride + kk] = hi;
        Bl[row * basis_stride + kk] = (_Float16)(v - (float)hi);
    }

    // generator role: column gc of the tile, bins 4 gh .. 4 gh + 3 of each 8-bin step
    const int gc = tid & (kProjCols - 1), gh = tid >> 7;
    const int gs = s0 + gc;
    auto xat = [&](int b, int part, int col) { return X + ((b * 2 + part) * kProjCols + col) * xs; };

    for (int a = a_begin; a < a_end; ++a) {
        __syncthreads();  // weight and X of the previous aperture are no longer read
        Column<P> col;
        col.lobe = -1;
        col.t = 0.5f;
        col.lobe_value = make_float2(0.f, 0.f);
#pragma unroll
        for (int r = 0; r < P; ++r) col.g[r] = make_float2(0.f, 0.f);
        if (gs < S) {
            const Pair q = pair_geometry(c, pos[3 * a], pos[3 * a + 1], pos[3 * a + 2], ctr[3 * gs], ctr[3 * gs + 1],
                                         ctr[3 * gs + 2], nrm[3 * gs], nrm[3 * gs + 1], nrm[3 * gs + 2]);
            col = column_of_pair<P>(c, q, make_float2(1.0f, 0.0f));
            if (gh == 0) weight[gc] = (q.amp * sqrtf(areas[gs])) * q.rot;
        } else if (gh == 0) {
            weight[gc] = make_float2(0.f, 0.f);
        }
        f32x16 acc[ROWS / 32];
#pragma unroll
        for (int x = 0; x < ROWS / 32; ++x)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[x][r] = 0.0f;
        // 4 bins x (re, im) of this thread's column for step t, as hi / lo, into X[t & 1]
        auto generate = [&](int t) {
            f16x8 hi, lo;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int k = t * 8 + gh * 4 + e;
                const float2 v = column_atom<P>(c, T, col, k, k % P);
                f16x2 h2, l2;
                split_f16_pair<false>(v.x, v.y, h2, l2);
                hi[2 * e] = h2.x, hi[2 * e + 1] = h2.y;
                lo[2 * e] = l2.x, lo[2 * e + 1] = l2.y;
            }
            *reinterpret_cast<f16x8*>(xat(t & 1, 0, gc) + 8 * gh) = hi;
            *reinterpret_cast<f16x8*>(xat(t & 1, 1, gc) + 8 * gh) = lo;
        };
        generate(0);
        __syncthreads();
        for (int t = 0; t < steps; ++t) {
            const int b = t & 1;
            const int colx = wave * 32 + l32;
            const f16x8 xh = *reinterpret_cast<const f16x8*>(xat(b, 0, colx) + 8 * h);
            const f16x8 xl = *reinterpret_cast<const f16x8*>(xat(b, 1, colx) + 8 * h);
#pragma unroll
            for (int x = 0; x < ROWS / 32; ++x) {
                const int row = x * 32 + l32;
                const f16x8 uh = *reinterpret_cast<const f16x8*>(Bh + row * basis_stride + t * kProjStep + 8 * h);
                const f16x8 ul = *reinterpret_cast<const f16x8*>(Bl + row * basis_stride + t * kProjStep + 8 * h);
                acc[x] = mfma_f16(uh, xh, acc[x]);
                acc[x] = mfma_f16(uh, xl, acc[x]);
                acc[x] = mfma_f16(ul, xh, acc[x]);
            }
            // the next step's atoms are generated while these products run on the matrix core;
            // X[b ^ 1] was last read in step t - 1, before the barrier that ended it
            if (t + 1 < steps) generate(t + 1);
            __syncthreads();
        }
        // (Re, Im) row pairs times the column weight; C[row][col] -> acc[r] with
        // row = 8 (r / 4) + 4 h + r % 4, so rows 2j and 2j + 1 are acc[r], acc[r + 1], r even
        const int colx = wave * 32 + l32, s = s0 + colx;
        const float2 w = weight[colx];
        if (s < S) {
#pragma unroll
            for (int x = 0; x < ROWS / 32; ++x)
#pragma unroll
                for (int r = 0; r < 16; r += 2) {
                    const int row = row0 + x * 32 + 8 * (r >> 2) + 4 * h + (r & 3);  // even
                    if (row < rows) {  // rows is even, so row + 1 < rows as well
                        const float hr = acc[x][r], hm = acc[x][r + 1];
                        out[((size_t)row * A + a) * S + s] = w.x * hr - w.y * hm;
                        out[((size_t)(row + 1) * A + a) * S + s] = w.x * hm + w.y * hr;
                    }
                }
        }
    }
}

static size_t project_lds(const Chirp& c, int rows_tile) {
    const size_t kk = (size_t)(c.K + 7) / 8 * kProjStep;
    return sizeof(float) * 2 * c.N + sizeof(float2) * kProjCols +
           sizeof(_Float16) * (2 * (size_t)rows_tile * (kk + 8) + 2 * 2 * kProjCols * (kProjStep + 8));
}

template <int P, int ROWS>
static hipError_t launch_project_pr(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                                    const float* areas, int S, const float* basis, int rows, float* out,
                                    hipStream_t st) {
    // ~2048 workgroups, but at least 16 apertures each so the basis load is amortized
    const int tiles = (S + kProjCols - 1) / kProjCols;
    int groups = (2048 + tiles - 1) / tiles;
    if (groups > (A + 15) / 16) groups = (A + 15) / 16;
    if (groups < 1) groups = 1;
    const int per_group = (A + groups - 1) / groups;
    groups = (A + per_group - 1) / per_group;
    const size_t smem = project_lds(c, ROWS);
    hipError_t e = allow_lds((const void*)project_kernel<P, ROWS>, smem);
    if (e != hipSuccess) return e;
    const int passes = (rows + ROWS - 1) / ROWS;
    hipLaunchKernelGGL((project_kernel<P, ROWS>), dim3(tiles, groups, passes), dim3(256), smem, st, c, pos, A, ctr,
                       nrm, areas, S, basis, rows, per_group, out);
    return hipGetLastError();
}

// The widest row tile (64 or 32 basis rows per pass) whose LDS fits; 0 if none does.
static int project_tile(const Chirp& c, int rows) {
    if (rows > 32 && project_lds(c, 64) <= 160 * 1024) return 64;
    return project_lds(c, 32) <= 160 * 1024 ? 32 : 0;
}

template <int P>
static hipError_t launch_project_p(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                                   const float* areas, int S, const float* basis, int rows, float* out,
                                   hipStream_t st) {
    if (project_tile(c, rows) == 64)
        return launch_project_pr<P, 64>(c, pos, A, ctr, nrm, areas, S, basis, rows, out, st);
    return launch_project_pr<P, 32>(c, pos, A, ctr, nrm, areas, S, basis, rows, out, st);
}

bool project_supported(const Chirp& c, int rows) { return rows >= 2 && rows % 2 == 0 && project_tile(c, rows) > 0; }

hipError_t launch_project(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                          const float* areas, int S, const float* basis, int rows, float* out, hipStream_t st) {
    switch (c.P) {
        case 1: return launch_project_p<1>(c, pos, A, ctr, nrm, areas, S, basis, rows, out, st);
        case 2: return launch_project_p<2>(c, pos, A, ctr, nrm, areas, S, basis, rows, out, st);
        case 4: return launch_project_p<4>(c, pos, A, ctr, nrm, areas, S, basis, rows, out, st);
        case 8: return launch_project_p<8>(c, pos, A, ctr, nrm, areas, S, basis, rows, out, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace dsplat
