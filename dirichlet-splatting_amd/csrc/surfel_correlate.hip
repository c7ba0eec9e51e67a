// Candidate correlations against a residual (reference: dsplat/kernels/correlate.slang:75-235;
// bindings dsplat/kernels/surfel.py:242-315).
//
// Layout (shared by every kernel here): a workgroup owns 32 apertures with their residual
// tile [K][32] in LDS; each half-wave takes one candidate and lane = aperture, reducing
// all bins of its pair in registers as per-class sums of the cot form
//   F_r = sum cot_k res_k,  Q_r = sum cot_k^2,  (gradient) G_r = sum cot_k^2 res_k,
//   C1_r = sum cot_k, C3_r = sum cot_k^3,  U_r = sum res_k, n_r = #bins of class r,
// from which, with the lobe bin taken out and added directly,
//   g0 = sum conj(D) res,  |D|^2 sum,  g1 = sum conj(D') res,  sum Re(conj(D') D).
// Each half-wave sums its per-aperture results over the 32 apertures and writes one
// partial per (candidate, aperture block); reduce_partials adds them in float64 in a
// fixed order. The reference adds them with float atomics instead (nondeterministic).
#include "dsplat_device.h"
#include "dsplat_internal.h"
#include "dsplat_sums.h"

namespace dsplat {

constexpr int kCorApertures = 32;

// Residual tile, class sums of the residual, and the derivative cot table of one workgroup.
struct CorTile {
    float* T;        // [2N], T_0 = 0
    float2* R;       // [Kp][32]
    float2* Us;      // [P][32]
};

__device__ __forceinline__ CorTile cor_tile(const Chirp& c, float* smem, int P) {
    CorTile t;
    t.T = smem;
    t.R = reinterpret_cast<float2*>(smem + 2 * c.N);
    t.Us = t.R + c.Kp * kCorApertures;
    return t;
}

__device__ __forceinline__ void load_tile(const Chirp& c, const CorTile& tile, int P, const float2* __restrict__ res,
                                          int A, int a0) {
    const int tid = threadIdx.x;
    for (int i = tid; i < 2 * c.N; i += blockDim.x) tile.T[i] = cot_entry(i, c.N, 0.0f);
    for (int i = tid; i < kCorApertures * c.Kp; i += blockDim.x) {
        const int ap = i / c.Kp, k = i - ap * c.Kp;
        const int a = a0 + ap;
        tile.R[k * kCorApertures + ap] = (a < A && k < c.K) ? res[(size_t)a * c.K + k] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    if (tid < P * kCorApertures) {
        const int r = tid / kCorApertures, ap = tid - r * kCorApertures;
        float2 s = make_float2(0.f, 0.f);
        for (int k = r; k < c.Kp; k += P) s += tile.R[k * kCorApertures + ap];
        tile.Us[r * kCorApertures + ap] = s;
    }
    __syncthreads();
}

// Bin sums of one pair (lane = aperture al of the tile). GRAD adds g1 and rx.
template <int P, bool GRAD>
__device__ __forceinline__ Sums pair_sums(const Chirp& c, const CorTile& tile, int al, const Pair& q) {
    return bin_sums<P, GRAD>(c, tile.T, tile.R + al, tile.Us + al, q, kCorApertures);
}

template <int NV>
__device__ __forceinline__ void half_wave_sum(float* v) {
#pragma unroll
    for (int o = 16; o >= 1; o >>= 1) {
#pragma unroll
        for (int j = 0; j < NV; ++j) v[j] += __shfl_xor(v[j], o, 64);
    }
}

// ---------------------------------------------------------------------------------------
// correlate (correlate.slang:78-108) and correlate_gradient (correlate.slang:174-235).
// Outputs per candidate (without area), in the gradient binding's column order:
//   [corr.re, corr.im, norm, dcorr.re[W], dcorr.im[W], dnorm[W]],  W = 0, 3 or 6.
// ---------------------------------------------------------------------------------------
template <int P, int W>
__global__ __launch_bounds__(256) void correlate_kernel(const Chirp c, const float* __restrict__ pos, int A,
                                                        const float* __restrict__ ctr,
                                                        const float* __restrict__ nrm, int S,
                                                        const float2* __restrict__ res, int chunk,
                                                        float* __restrict__ partial) {
    constexpr int NV = 3 + 3 * W;
    extern __shared__ float smem[];
    const CorTile tile = cor_tile(c, smem, P);
    const int blk = blockIdx.x, nblk = gridDim.x, a0 = blk * kCorApertures;
    load_tile(c, tile, P, res, A, a0);
    const int tid = threadIdx.x, wave = tid >> 6, half = (tid >> 5) & 1, al = tid & 31;
    const int a = min(a0 + al, A - 1);
    const bool aperture_ok = a0 + al < A;
    const float px = pos[3 * a], py = pos[3 * a + 1], pz = pos[3 * a + 2];
    const int s_begin = blockIdx.y * chunk, s_end = min(S, s_begin + chunk);
    for (int s = s_begin + 2 * wave + half; s < s_end; s += 8) {
        const float nx = nrm[3 * s], ny = nrm[3 * s + 1], nz = nrm[3 * s + 2];
        Pair q = pair_geometry(c, px, py, pz, ctr[3 * s], ctr[3 * s + 1], ctr[3 * s + 2], nx, ny, nz);
        if (!aperture_ok) { q.amp = 0.0f; q.amp_cos = 0.0f; }
        const Sums z = pair_sums<P, (W > 0)>(c, tile, al, q);
        const float2 corr = cmulc(q.rot, z.g0);  // sum conj(D rot) res
        float v[NV];
        v[0] = q.amp * corr.x;
        v[1] = q.amp * corr.y;
        v[2] = q.amp * q.amp * z.nrm;
        if (W > 0) {
            const float2 corr_x = cmulc(q.rot, z.g1);
            float da[3];
            amp_gradient(q, nx, ny, nz, da);
            const float u[3] = {q.ux, q.uy, q.uz};
#pragma unroll
            for (int j = 0; j < W; ++j) {
                float2 dc;
                float dn;
                if (j < 3) {
                    const float2 icorr = make_float2(-corr.y, corr.x);
                    dc = da[j] * corr + (q.amp * u[j]) * (c.k0pm_f * corr_x + q.rate * icorr);
                    dn = 2.0f * q.amp * (da[j] * z.nrm + q.amp * c.k0pm_f * u[j] * z.rx);
                } else {
                    const float d = q.amp_cos * u[j - 3];
                    dc = d * corr;
                    dn = 2.0f * q.amp * d * z.nrm;
                }
                v[3 + j] = dc.x;
                v[3 + W + j] = dc.y;
                v[3 + 2 * W + j] = dn;
            }
        }
        half_wave_sum<NV>(v);
        if (al == 0) {
            float* dst = partial + ((size_t)s * nblk + blk) * NV;
#pragma unroll
            for (int j = 0; j < NV; ++j) dst[j] = v[j];
        }
    }
}

// ---------------------------------------------------------------------------------------
// correlate_orientations (correlate.slang:110-169): distance, phase and the bin sums are
// shared by the own normal (j = 0) and B bank normals; only the amplitude changes.
// With a range table the bin sum is replaced by a 4-tap Catmull-Rom interpolation of
// the FFT-oversampled correlation of each aperture (fmcw.py:337-355).
// Outputs per candidate: [1 + B][corr.re, corr.im, norm].
// ---------------------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(256) void orientations_kernel(
    const Chirp c, const float* __restrict__ pos, int A, const float* __restrict__ ctr, const float* __restrict__ nrm,
    const float* __restrict__ bank, int B, int S, const float2* __restrict__ res, const float2* __restrict__ table,
    const float* __restrict__ norm_table, int oversample, int chunk, float* __restrict__ partial) {
    extern __shared__ float smem[];
    const CorTile tile = cor_tile(c, smem, P);
    const int blk = blockIdx.x, nblk = gridDim.x, a0 = blk * kCorApertures;
    if (!table) load_tile(c, tile, P, res, A, a0);
    const int tid = threadIdx.x, wave = tid >> 6, half = (tid >> 5) & 1, al = tid & 31;
    const int a = min(a0 + al, A - 1);
    const bool aperture_ok = a0 + al < A;
    const float px = pos[3 * a], py = pos[3 * a + 1], pz = pos[3 * a + 2];
    const int s_begin = blockIdx.y * chunk, s_end = min(S, s_begin + chunk);
    const int NV = 3 * (1 + B);
    for (int s = s_begin + 2 * wave + half; s < s_end; s += 8) {
        const float nx = nrm[3 * s], ny = nrm[3 * s + 1], nz = nrm[3 * s + 2];
        const Pair q = pair_geometry(c, px, py, pz, ctr[3 * s], ctr[3 * s + 1], ctr[3 * s + 2], nx, ny, nz);
        float2 corr;
        float zn;
        if (table) {
            const int size = c.N * oversample;
            const double x = (double)q.r * c.k0pm * oversample;
            const double fl = floor(x);
            const float tt = (float)(x - fl);
            const int base = (int)fl;
            const float t2 = tt * tt, t3 = t2 * tt;
            const float wts[4] = {-0.5f * tt + t2 - 0.5f * t3, 1.0f - 2.5f * t2 + 1.5f * t3,
                                  0.5f * tt + 2.0f * t2 - 1.5f * t3, -0.5f * t2 + 0.5f * t3};
            float2 value = make_float2(0.f, 0.f);
            zn = 0.0f;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                int k = (base + j - 1) % size;
                if (k < 0) k += size;
                value += wts[j] * table[(size_t)a * size + k];
                zn = fmaf(wts[j], norm_table[k], zn);
            }
            corr = cmul(value, make_float2(q.rot.x, -q.rot.y));
            zn = fmaxf(zn, 0.0f);
        } else {
            const Sums z = pair_sums<P, false>(c, tile, al, q);
            corr = cmulc(q.rot, z.g0);
            zn = z.nrm;
        }
        float* dst = partial + ((size_t)s * nblk + blk) * NV;
        for (int j = 0; j <= B; ++j) {
            const float mx = j == 0 ? nx : bank[3 * (j - 1)];
            const float my = j == 0 ? ny : bank[3 * (j - 1) + 1];
            const float mz = j == 0 ? nz : bank[3 * (j - 1) + 2];
            const float cosine = q.ux * mx + q.uy * my + q.uz * mz;
            float amp = 0.0f;
            if (cosine > 0.0f && aperture_ok) {
                float f = 1.0f;
                if (c.fresnel_n > 1.0f) {
                    const float root = sqrtf(c.fresnel_n * c.fresnel_n - 1.0f + cosine * cosine);
                    f = (root - cosine) / (root + cosine);
                }
                amp = c.lam * q.inv_r * (1.0f / (8.0f * kPi)) * cosine * cosine * f;
            }
            float v[3] = {amp * corr.x, amp * corr.y, amp * amp * zn};
            half_wave_sum<3>(v);
            if (al == 0) {
                dst[3 * j] = v[0];
                dst[3 * j + 1] = v[1];
                dst[3 * j + 2] = v[2];
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// Batched certificate scan: every scan of a set (uneven aperture counts) in one launch.
// blockIdx.x runs over the aperture blocks of all scans back to back; a workgroup finds
// its scan in the set's prefix table, loads that scan's residual tile and correlates a
// chunk of candidates (blockIdx.y) exactly as correlate_kernel<P, 0>.
// partial [C][total blocks][3]; scan_reduce sums each scan's block range in float64.
// ---------------------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(256) void scan_kernel(const Chirp c, const ScanSet set, const float* __restrict__ ctr,
                                                   const float* __restrict__ nrm, int S, int chunk,
                                                   float* __restrict__ partial) {
    extern __shared__ float smem[];
    const CorTile tile = cor_tile(c, smem, P);
    const int gblk = blockIdx.x, nblk = gridDim.x;
    int sc = 0;
    while (sc + 1 < set.count && gblk >= set.first_block[sc + 1]) ++sc;
    const int A = set.apertures[sc];
    const float* __restrict__ pos = set.positions[sc];
    const int a0 = (gblk - set.first_block[sc]) * kCorApertures;
    load_tile(c, tile, P, set.residuals[sc], A, a0);
    const int tid = threadIdx.x, wave = tid >> 6, half = (tid >> 5) & 1, al = tid & 31;
    const int a = min(a0 + al, A - 1);
    const bool aperture_ok = a0 + al < A;
    const float px = pos[3 * a], py = pos[3 * a + 1], pz = pos[3 * a + 2];
    const int s_begin = blockIdx.y * chunk, s_end = min(S, s_begin + chunk);
    for (int s = s_begin + 2 * wave + half; s < s_end; s += 8) {
        Pair q = pair_geometry(c, px, py, pz, ctr[3 * s], ctr[3 * s + 1], ctr[3 * s + 2], nrm[3 * s], nrm[3 * s + 1],
                               nrm[3 * s + 2]);
        if (!aperture_ok) q.amp = 0.0f;
        const Sums z = pair_sums<P, false>(c, tile, al, q);
        const float2 corr = cmulc(q.rot, z.g0);
        float v[3] = {q.amp * corr.x, q.amp * corr.y, q.amp * q.amp * z.nrm};
        half_wave_sum<3>(v);
        if (al == 0) {
            float* dst = partial + ((size_t)s * nblk + gblk) * 3;
            dst[0] = v[0];
            dst[1] = v[1];
            dst[2] = v[2];
        }
    }
}

// out[scan][s][3] = sum of partial[s][b][3] over the scan's blocks, float64, fixed order.
__global__ __launch_bounds__(256) void scan_reduce(const float* __restrict__ partial, int S, int nblk,
                                                   const ScanSet set, float* __restrict__ out) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int sc = blockIdx.y;
    if (row >= S) return;
    const int b0 = set.first_block[sc], b1 = set.first_block[sc + 1];
    for (int f = 0; f < 3; ++f) {
        double acc = 0.0;
        for (int b = b0 + lane; b < b1; b += 64) acc += partial[((size_t)row * nblk + b) * 3 + f];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (lane == 0) out[((size_t)sc * S + row) * 3 + f] = (float)acc;
    }
}

template <int P>
static hipError_t launch_scan_p(const Chirp& c, const ScanSet& set, const float* ctr, const float* nrm, int S,
                                int chunk, float* partial, float* out, hipStream_t st) {
    const int nblk = set.first_block[set.count];
    const size_t smem = cor_smem(c);
    hipError_t e = allow_lds((const void*)scan_kernel<P>, smem);
    if (e != hipSuccess) return e;
    const int groups = (S + chunk - 1) / chunk;
    hipLaunchKernelGGL((scan_kernel<P>), dim3(nblk, groups), dim3(256), smem, st, c, set, ctr, nrm, S, chunk, partial);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(scan_reduce, dim3((S + 3) / 4, set.count), dim3(256), 0, st, partial, S, nblk, set, out);
    return hipGetLastError();
}

hipError_t launch_scan(const Chirp& c, const ScanSet& set, const float* ctr, const float* nrm, int S, int chunk,
                       float* partial, float* out, hipStream_t st) {
    switch (c.P) {
        case 1: return launch_scan_p<1>(c, set, ctr, nrm, S, chunk, partial, out, st);
        case 2: return launch_scan_p<2>(c, set, ctr, nrm, S, chunk, partial, out, st);
        case 4: return launch_scan_p<4>(c, set, ctr, nrm, S, chunk, partial, out, st);
        case 8: return launch_scan_p<8>(c, set, ctr, nrm, S, chunk, partial, out, st);
        default: return hipErrorInvalidValue;
    }
}

// out[s, f] = sum over blocks of partial[s, blk, f] in float64, fixed order (one wave per row).
__global__ __launch_bounds__(256) void reduce_partials(const float* __restrict__ partial, int rows, int nblk,
                                                       int width, float* __restrict__ out) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= rows) return;
    for (int f = 0; f < width; ++f) {
        double acc = 0.0;
        for (int b = lane; b < nblk; b += 64) acc += partial[((size_t)row * nblk + b) * width + f];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (lane == 0) out[(size_t)row * width + f] = (float)acc;
    }
}

hipError_t launch_reduce_partials(const float* partial, int rows, int nblk, int width, float* out, hipStream_t st) {
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(reduce_partials, dim3((rows + 3) / 4), dim3(256), 0, st, partial, rows, nblk, width, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
CorShape cor_shape(int A, int S, int target) {
    CorShape v;
    v.nblk = (A + kCorApertures - 1) / kCorApertures;
    int groups = (target + v.nblk - 1) / v.nblk;
    const int max_groups = (S + 7) / 8;
    groups = groups < 1 ? 1 : (groups > max_groups ? max_groups : groups);
    const int chunk = (S + groups - 1) / groups;
    v.chunk = (chunk + 7) / 8 * 8;
    v.groups = (S + v.chunk - 1) / v.chunk;
    return v;
}

size_t cor_smem(const Chirp& c) {
    return sizeof(float) * 2 * c.N + sizeof(float2) * ((size_t)c.Kp * kCorApertures + 8 * kCorApertures);
}

template <int P, int W>
static hipError_t launch_cor_pw(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm, int S,
                                const float2* res, float* partial, float* out, hipStream_t st) {
    // up to a few thousand candidates: one round of workgroups (two per CU: ~70 KB of LDS
    // each), candidates split only as far as that needs, so each workgroup loads its residual
    // tile once for many candidates (measured 1.5x faster at 64 candidates, 3 % at 2035);
    // larger candidate sets keep eight rounds (measured 1.7x faster at 262144)
    const CorShape v = cor_shape(A, S, S <= kCorFewCandidates ? kCorRound : 4096);
    const size_t smem = cor_smem(c);
    hipError_t e = allow_lds((const void*)correlate_kernel<P, W>, smem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((correlate_kernel<P, W>), dim3(v.nblk, v.groups), dim3(256), smem, st, c, pos, A, ctr, nrm, S,
                       res, v.chunk, partial);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(reduce_partials, dim3((S + 3) / 4), dim3(256), 0, st, partial, S, v.nblk, 3 + 3 * W, out);
    return hipGetLastError();
}

template <int P>
static hipError_t launch_cor_p(const Chirp& c, int W, const float* pos, int A, const float* ctr, const float* nrm,
                               int S, const float2* res, float* partial, float* out, hipStream_t st) {
    switch (W) {
        case 0: return launch_cor_pw<P, 0>(c, pos, A, ctr, nrm, S, res, partial, out, st);
        case 3: return launch_cor_pw<P, 3>(c, pos, A, ctr, nrm, S, res, partial, out, st);
        case 6: return launch_cor_pw<P, 6>(c, pos, A, ctr, nrm, S, res, partial, out, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_correlate(const Chirp& c, int W, const float* pos, int A, const float* ctr, const float* nrm, int S,
                            const float2* res, float* partial, float* out, hipStream_t st) {
    switch (c.P) {
        case 1: return launch_cor_p<1>(c, W, pos, A, ctr, nrm, S, res, partial, out, st);
        case 2: return launch_cor_p<2>(c, W, pos, A, ctr, nrm, S, res, partial, out, st);
        case 4: return launch_cor_p<4>(c, W, pos, A, ctr, nrm, S, res, partial, out, st);
        case 8: return launch_cor_p<8>(c, W, pos, A, ctr, nrm, S, res, partial, out, st);
        default: return hipErrorInvalidValue;
    }
}

// Best orientation per candidate in one pass (lane = candidate): every aperture's pair
// geometry and bin sum (or 4-tap range-table interpolation) is shared by the own normal
// and the B bank normals, whose sums stay in registers; after the last aperture the
// certificates |s|^2 / max(area |a|^2, 1e-60) (s = sqrt(area) a^H r, or its real part)
// are formed in float64 and the first maximum is kept (NaN counts as largest). No
// per-block partials. The aperture sums accumulate in float32 in aperture order.
constexpr int kOrientMax = 25;  // own normal + up to 24 bank normals

template <int P, bool TABLE, bool FRESNEL>
__global__ __launch_bounds__(256) void orient_best_kernel(
    const Chirp c, const float* __restrict__ pos, int A, const float* __restrict__ ctr, const float* __restrict__ nrm,
    const float* __restrict__ areas, const float* __restrict__ bank, int B, int S, const float2* __restrict__ res,
    const float2* __restrict__ table, const float* __restrict__ norm_table, int oversample, int real_part,
    double* __restrict__ best_value, int* __restrict__ best_index) {
    extern __shared__ float smem[];
    __shared__ float dirs[3 * kOrientMax];
    const int tid = threadIdx.x;
    const int s = blockIdx.x * 256 + tid;
    const int sc = s < S ? s : S - 1;
    const float cx = ctr[3 * sc], cy = ctr[3 * sc + 1], cz = ctr[3 * sc + 2];
    const float nx = nrm[3 * sc], ny = nrm[3 * sc + 1], nz = nrm[3 * sc + 2];
    for (int i = tid; i < 3 * B; i += 256) dirs[3 + i] = bank[i];
    float acc[kOrientMax][3];
#pragma unroll
    for (int j = 0; j < kOrientMax; ++j) acc[j][0] = acc[j][1] = acc[j][2] = 0.0f;
    for (int i = 3 * (B + 1) + tid; i < 3 * kOrientMax; i += 256) dirs[i] = 0.0f;  // unused slots: zero normals
    // amp = lam / (8 pi r) cos^2 F(cos) for cos > 0, else 0 (scattering.slang:16-38), with the
    // Fresnel factor F = (root - cos) / (root + cos) = (root - cos)^2 / (n^2 - 1),
    // root = sqrt(n^2 - 1 + cos^2): no division per normal, and cos <= 0 enters as 0.
    const float n2m1 = FRESNEL ? c.fresnel_n * c.fresnel_n - 1.0f : 1.0f;
    const float amp_scale = c.lam * (1.0f / (8.0f * kPi)) / n2m1;
    auto accumulate = [&](const Pair& q, float2 corr, float zn) {
        const float pre = amp_scale * q.inv_r;
        // every slot is evaluated (a zero normal adds nothing), so the sums stay in registers
#pragma unroll
        for (int j = 0; j < kOrientMax; ++j) {
            const float mx = j == 0 ? nx : dirs[3 * j], my = j == 0 ? ny : dirs[3 * j + 1];
            const float mz = j == 0 ? nz : dirs[3 * j + 2];
            const float cp = fmaxf(q.ux * mx + q.uy * my + q.uz * mz, 0.0f);
            float amp = pre * cp * cp;
            if constexpr (FRESNEL) {
                const float d = __builtin_amdgcn_sqrtf(fmaf(cp, cp, n2m1)) - cp;  // v_sqrt_f32, 1 ulp
                amp *= d * d;
            }
            acc[j][0] = fmaf(amp, corr.x, acc[j][0]);
            acc[j][1] = fmaf(amp, corr.y, acc[j][1]);
            acc[j][2] = fmaf(amp * amp, zn, acc[j][2]);
        }
    };
    __syncthreads();
    if constexpr (TABLE) {
        const int size = c.N * oversample;
        for (int a = 0; a < A; ++a) {
            const Pair q =
                pair_geometry_t<FRESNEL>(c, pos[3 * a], pos[3 * a + 1], pos[3 * a + 2], cx, cy, cz, nx, ny, nz);
            const double x = (double)q.r * c.k0pm * oversample;
            const double fl = floor(x);
            const float tt = (float)(x - fl);
            const int base = (int)fl;
            const float t2 = tt * tt, t3 = t2 * tt;
            const float wts[4] = {-0.5f * tt + t2 - 0.5f * t3, 1.0f - 2.5f * t2 + 1.5f * t3,
                                  0.5f * tt + 2.0f * t2 - 1.5f * t3, -0.5f * t2 + 0.5f * t3};
            float2 value = make_float2(0.f, 0.f);
            float zn = 0.0f;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                int k = (base + j - 1) % size;
                if (k < 0) k += size;
                value += wts[j] * table[(size_t)a * size + k];
                zn = fmaf(wts[j], norm_table[k], zn);
            }
            accumulate(q, cmul(value, make_float2(q.rot.x, -q.rot.y)), fmaxf(zn, 0.0f));
        }
    } else {
        const CorTile tile = cor_tile(c, smem, P);
        for (int a0 = 0; a0 < A; a0 += kCorApertures) {
            __syncthreads();  // the previous tile is no longer read
            load_tile(c, tile, P, res, A, a0);
            const int na = min(kCorApertures, A - a0);
            for (int al = 0; al < na; ++al) {
                const int a = a0 + al;
                const Pair q =
                    pair_geometry_t<FRESNEL>(c, pos[3 * a], pos[3 * a + 1], pos[3 * a + 2], cx, cy, cz, nx, ny, nz);
                const Sums z = pair_sums<P, false>(c, tile, al, q);
                accumulate(q, cmulc(q.rot, z.g0), z.nrm);
            }
        }
    }
    if (s >= S) return;
    const float scale = sqrtf(areas[s]);
    const float scale2 = scale * scale;
    double best = 0.0;
    int index = 0;
#pragma unroll
    for (int j = 0; j < kOrientMax; ++j) {
        if (j > B) continue;
        const double re = (double)(acc[j][0] * scale), im = (double)(acc[j][1] * scale);
        const double mag = real_part ? fabs(re) : hypot(re, im);
        const double value = mag * mag / fmax((double)(acc[j][2] * scale2), 1e-60);
        if (j == 0 || (isnan(value) && !isnan(best)) || (!isnan(best) && value > best)) {
            best = value;
            index = j;
        }
    }
    best_value[s] = best;
    best_index[s] = index;
}

template <int P, bool F>
static hipError_t launch_orient_best_pf(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                                        const float* bank, int B, int S, const float2* res, const float2* table,
                                        const float* norm_table, int oversample, const OrientPick& pick,
                                        hipStream_t st) {
    const dim3 grid((S + 255) / 256);
    hipError_t e;
    if (table) {
        hipLaunchKernelGGL((orient_best_kernel<P, true, F>), grid, dim3(256), 0, st, c, pos, A, ctr, nrm, pick.areas,
                           bank, B, S, res, table, norm_table, oversample, pick.real_part, pick.value, pick.index);
    } else {
        const size_t smem = cor_smem(c);
        if ((e = allow_lds((const void*)orient_best_kernel<P, false, F>, smem)) != hipSuccess) return e;
        hipLaunchKernelGGL((orient_best_kernel<P, false, F>), grid, dim3(256), smem, st, c, pos, A, ctr, nrm,
                           pick.areas, bank, B, S, res, table, norm_table, oversample, pick.real_part, pick.value,
                           pick.index);
    }
    return hipGetLastError();
}

template <int P>
static hipError_t launch_orient_best_p(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                                       const float* bank, int B, int S, const float2* res, const float2* table,
                                       const float* norm_table, int oversample, const OrientPick& pick,
                                       hipStream_t st) {
    if (c.fresnel_n > 1.0f)
        return launch_orient_best_pf<P, true>(c, pos, A, ctr, nrm, bank, B, S, res, table, norm_table, oversample,
                                              pick, st);
    return launch_orient_best_pf<P, false>(c, pos, A, ctr, nrm, bank, B, S, res, table, norm_table, oversample, pick,
                                           st);
}

hipError_t launch_orient_best(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                              const float* bank, int B, int S, const float2* res, const float2* table,
                              const float* norm_table, int oversample, const OrientPick& pick, hipStream_t st) {
    if (B + 1 > kOrientMax) return hipErrorInvalidValue;
    switch (c.P) {
        case 1: return launch_orient_best_p<1>(c, pos, A, ctr, nrm, bank, B, S, res, table, norm_table, oversample, pick, st);
        case 2: return launch_orient_best_p<2>(c, pos, A, ctr, nrm, bank, B, S, res, table, norm_table, oversample, pick, st);
        case 4: return launch_orient_best_p<4>(c, pos, A, ctr, nrm, bank, B, S, res, table, norm_table, oversample, pick, st);
        case 8: return launch_orient_best_p<8>(c, pos, A, ctr, nrm, bank, B, S, res, table, norm_table, oversample, pick, st);
        default: return hipErrorInvalidValue;
    }
}

template <int P>
static hipError_t launch_orient_p(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                                  const float* bank, int B, int S, const float2* res, const float2* table,
                                  const float* norm_table, int oversample, float* partial, float* out,
                                  hipStream_t st) {
    const CorShape v = cor_shape(A, S);
    const size_t smem = table ? 0 : cor_smem(c);
    hipError_t e = allow_lds((const void*)orientations_kernel<P>, smem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((orientations_kernel<P>), dim3(v.nblk, v.groups), dim3(256), smem, st, c, pos, A, ctr, nrm, bank,
                       B, S, res, table, norm_table, oversample, v.chunk, partial);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(reduce_partials, dim3((S + 3) / 4), dim3(256), 0, st, partial, S, v.nblk, 3 * (1 + B), out);
    return hipGetLastError();
}

hipError_t launch_orientations(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                               const float* bank, int B, int S, const float2* res, const float2* table,
                               const float* norm_table, int oversample, float* partial, float* out,
                               hipStream_t st) {
    switch (c.P) {
        case 1: return launch_orient_p<1>(c, pos, A, ctr, nrm, bank, B, S, res, table, norm_table, oversample, partial, out, st);
        case 2: return launch_orient_p<2>(c, pos, A, ctr, nrm, bank, B, S, res, table, norm_table, oversample, partial, out, st);
        case 4: return launch_orient_p<4>(c, pos, A, ctr, nrm, bank, B, S, res, table, norm_table, oversample, partial, out, st);
        case 8: return launch_orient_p<8>(c, pos, A, ctr, nrm, bank, B, S, res, table, norm_table, oversample, partial, out, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace dsplat
