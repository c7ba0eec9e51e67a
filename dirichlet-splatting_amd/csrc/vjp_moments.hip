// Surfel VJP in the sample domain by bin moments (reference semantics:
// dsplat/kernels/surfel.slang:257-311, binding dsplat/kernels/surfel.py:123-137).
//
// Per pair (a, s) the VJP needs
//   G0 = sum_n w^n U[a, n],   G1 = sum_n n w^n U[a, n],   w = e^{-2 pi i k0 / N},
// with U[a, n] = sum_{k<K} e^{2 pi i n k / N} u[a, k] the cotangent in the sample domain and
// k0 = lobe + delta (|delta| <= 1/2) the pair's peak bin. With n0 = (S - 1) / 2, h = S / 2,
// x_n = (n - n0) / h in [-1, 1] and t = pi delta S / N (|t| <= pi / 4 for pad >= 2),
//   w^n = e^{-2 pi i lobe n / N} rho e^{-i t x_n},   rho = e^{-2 pi i delta n0 / N},
// and the Taylor series of e^{-i t x} gives
//   G0 = rho P(t),   G1 = n0 G0 + i h rho P'(t),   P(t) = sum_{q<Q} M_q[lobe] t^q,
//   M_q[m] = ((-i)^q / q!) sum_n x_n^q e^{-2 pi i m n / N} U[a, n].
// Q = 10 moments truncate at (pi / 4)^10 / 10! = 2.5e-8 relative. So a pair costs one row of
// 10 complex moments (80 bytes of LDS), a 10-step Horner evaluation of P and P' and one small
// rotation, besides its geometry: no per-pair powers, no operand splits, no matrix products.
//
// The moments depend on the aperture and the bin only. vjp_moments_prep forms them for the
// band of bins an aperture's pairs can reach: the lobes of all pairs of aperture a lie
// within one bin of the range interval from a to the box of the centres, and the band starts
// one bin below it (m0, span <= 32 rows). An aperture whose interval is wider (a scene more
// than ~30 bins deep, seen from that aperture) keeps U in its slot instead, and its pairs
// take the exact sum over the samples: slower, never wrong.
//
// vjp_moments_kernel: a workgroup owns 256 surfels (one per lane) and a range of apertures,
// whose slots are staged 8 at a time by direct global -> LDS loads. The 8 chained gradients
// accumulate in registers over the range; one partial per (surfel, range) is summed in
// float64 by vjp_reduce. Workgroup order is XCD-aware: the surfel tiles of one range run on
// one XCD and read its slots through one L2.
#include "dsplat_device.h"
#include "dsplat_internal.h"
#include "dsplat_td.h"

namespace dsplat {

constexpr int kMomQ = 10;                   // Taylor moments per bin
constexpr int kMomBand = 32;                // bins per aperture band
constexpr int kMomSlot = kMomBand * kMomQ;  // float2 per aperture slot (2560 bytes)
constexpr int kMomGroup = 8;                // apertures staged per LDS fill
constexpr int kMomSurfels = 256;            // surfels per workgroup
constexpr int kMomSlotU4 = kMomSlot / 2;    // uint4 per slot
static_assert(kMomSlot >= 256, "a slot holds the 256 samples of a wide aperture");
static_assert((kMomGroup * kMomSlotU4) % 256 == 0, "whole staging rounds");

bool vjp_moments_supported(const Chirp& c) {
    // |t| <= pi S / (2 N) <= pi / 4 needs pad >= 2; the pre-pass DFT takes 16 rows of K bins
    return (c.S == 128 || c.S == 256) && c.N >= 2 * c.S && (c.N & (c.N - 1)) == 0 && c.K % 8 == 0;
}

struct MomShape {
    int tiles, ranges, per_range;
};

static MomShape mom_shape(int A, int S) {
    MomShape v;
    v.tiles = (S + kMomSurfels - 1) / kMomSurfels;
    const int max_ranges = (A + kMomGroup - 1) / kMomGroup;
    int ranges = (2048 + v.tiles - 1) / v.tiles;  // about 8 workgroups per CU
    ranges = ranges < 1 ? 1 : (ranges > max_ranges ? max_ranges : ranges);
    int per = (A + ranges - 1) / ranges;
    v.per_range = (per + kMomGroup - 1) / kMomGroup * kMomGroup;
    v.ranges = (A + v.per_range - 1) / v.per_range;
    return v;
}

// Workspace: partials [S][ranges][8] floats | slots [A][kMomSlot] float2 | bands [A] int2,
// each part 256-byte aligned.
struct MomLayout {
    size_t partial, slots, bands, total;
};

static MomLayout mom_layout(int A, int S) {
    const MomShape v = mom_shape(A, S);
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    MomLayout w;
    w.partial = 0;
    w.slots = up(sizeof(float) * 8 * (size_t)S * v.ranges);
    w.bands = w.slots + up(sizeof(float2) * kMomSlot * (size_t)A);
    w.total = w.bands + up(sizeof(int2) * (size_t)A);
    return w;
}

size_t vjp_moments_workspace(int A, int S) { return mom_layout(A, S).total; }

// Apertures per pre-pass workgroup: the FFT buffer holds G rows of N points (32 KB).
__host__ __device__ constexpr int mom_prep_rows(int N) { return N <= 256 ? 16 : 4096 / N; }

// G apertures per workgroup (mom_prep_rows). The cotangent rows go to the sample domain by a
// radix-2 FFT in LDS: U[n] = sum_{k<K} e^{2 pi i n k / N} u[k] for n < N, in bit-reversed order
// (decimation in frequency; u is zero past K = N / 2, which the first stage folds in). Each
// aperture's band comes from the box of the centres, and the moments of its band rows from a
// small product on the matrix cores (below); a wide aperture's slot gets U.
__global__ __launch_bounds__(256) void vjp_moments_prep(const Chirp c, const float* __restrict__ pos, int A,
                                                        const float* __restrict__ ctr, int S,
                                                        const float2* __restrict__ up, float2* __restrict__ slots,
                                                        int2* __restrict__ bands) {
    extern __shared__ float smem[];
    const int G = mom_prep_rows(c.N), NH = c.N >> 1;
    float2* E = reinterpret_cast<float2*>(smem);  // [N]
    float2* F = E + c.N;                          // [G][N], U in bit-reversed order
    float* X = reinterpret_cast<float*>(F + G * c.N);  // [S][10]: x_n^q / q!
    __shared__ float box[4][6];
    __shared__ int2 bd[16];
    const int tid = threadIdx.x, wave = tid >> 6, L = tid & 63;
    const int r0 = blockIdx.x * G;
    const int bits = 31 - __builtin_clz(c.N);  // log2 N
    fill_twiddles(E, c.N, tid, 256);
    // box of the centres: lower (0..2) and upper (3..5) corner
    float bx[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int s = tid; s < S; s += 256) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float x = ctr[3 * s + d];
            bx[d] = fminf(bx[d], x);
            bx[3 + d] = fmaxf(bx[3 + d], x);
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            bx[d] = fminf(bx[d], __shfl_xor(bx[d], o, 64));
            bx[3 + d] = fmaxf(bx[3 + d], __shfl_xor(bx[3 + d], o, 64));
        }
    }
    if (L == 0) {
#pragma unroll
        for (int d = 0; d < 6; ++d) box[wave][d] = bx[d];
    }
    // powers X[n][q] = x_n^q / q! (q < 10; the product's columns 10..15 are zero)
    for (int n = tid; n < c.S; n += 256) {
        const float x = ((float)n - 0.5f * (float)(c.S - 1)) * (2.0f / (float)c.S);
        float pw = 1.0f;
#pragma unroll
        for (int q = 0; q < kMomQ; ++q) {
            X[n * kMomQ + q] = pw;
            pw = pw * x / (float)(q + 1);
        }
    }
    __syncthreads();  // twiddles written
    // first FFT stage with the zero upper half: F[k] = u[k], F[k + N / 2] = u[k] e^{2 pi i k / N}
    for (int i = tid; i < G * NH; i += 256) {
        const int r = i / NH, k = i - r * NH;
        const float2 u = r0 + r < A ? up[(size_t)(r0 + r) * c.K + k] : make_float2(0.f, 0.f);
        float2* row = F + r * c.N;
        row[k] = u;
        row[k + NH] = cmul(u, E[k]);
    }
    __syncthreads();
    for (int half = c.N >> 2; half >= 1; half >>= 1) {
        const int tw = NH / half;  // twiddle index step: e^{2 pi i j / (2 half)} = E[j N / (2 half)]
        for (int i = tid; i < G * NH; i += 256) {
            const int r = i / NH, t = i - r * NH;
            const int j = t & (half - 1), at = 2 * (t - j) + j;
            float2* row = F + r * c.N;
            const float2 a = row[at], b = row[at + half];
            row[at] = a + b;
            row[at + half] = cmul(a - b, E[j * tw]);
        }
        __syncthreads();
    }
    if (tid < G) {
        // distance interval from the aperture to the box, in float64; lobes lie within one
        // bin of [rint(k_lo), rint(k_hi)], so the band starts at rint(k_lo) - 1
        const int a = min(r0 + tid, A - 1);
        double near2 = 0.0, far2 = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double lo = fmin(fmin(box[0][d], box[1][d]), fmin(box[2][d], box[3][d]));
            const double hi = fmax(fmax(box[0][3 + d], box[1][3 + d]), fmax(box[2][3 + d], box[3][3 + d]));
            const double p = pos[3 * a + d];
            const double out = fmax(fmax(lo - p, p - hi), 0.0);
            const double span = fmax(fabs(p - lo), fabs(p - hi));
            near2 += out * out;
            far2 += span * span;
        }
        const double m_lo = rint(sqrt(near2) * c.k0pm) - 1.0, m_hi = rint(sqrt(far2) * c.k0pm) + 1.0;
        const double span = m_hi - m_lo + 1.0;
        const double m0 = m_lo - (double)c.N * floor(m_lo / (double)c.N);
        bd[tid] = make_int2((int)m0 & (c.N - 1), span <= (double)kMomBand ? (int)span : 0);  // 0: wide
        if (r0 + tid < A) bands[r0 + tid] = bd[tid];
    }
    __syncthreads();
    // Per aperture the moments are a real product on v_mfma_f32_16x16x4f32 (A[i][k] -> lane
    // i + 16 k, B[k][q] -> lane q + 16 k, C[4 (lane / 16) + r][lane % 16] in register r):
    // rows i = (j, part) of 8 band rows j per tile, K = samples n = 4 t + k,
    //   A[(j, part)][n] = Re / Im of e^{-2 pi i m_j n / N} U[n],  B[n][q] = x_n^q / q!,
    // and C[(j, part)][q] = the real or imaginary part of M_q[m_j] before (-i)^q. Each wave takes
    // every fourth aperture of the group.
    const int kk = L >> 4, col = L & 15;
    for (int qa = wave; qa < G; qa += 4) {
        const int a = r0 + qa;
        if (a >= A) break;  // uniform over the wave
        const int2 b = bd[qa];
        const float2* Fa = F + qa * c.N;  // U[n] = Fa[rev(n)]
        float2* slot = slots + (size_t)a * kMomSlot;
        if (b.y == 0) {
            for (int n = L; n < c.S; n += 64) slot[n] = Fa[__builtin_bitreverse32(n) >> (32 - bits)];
            continue;
        }
        for (int T = 0; 8 * T < b.y; ++T) {
            const int j = 8 * T + (col >> 1), part = col & 1;  // this lane's A row
            const int m = (b.x + j) & (c.N - 1);
            const int step = (4 * m) & (c.N - 1);
            int idx = (m * kk) & (c.N - 1);  // m n mod N at n = 4 t + kk
            // the reads of 8 steps go ahead of their products; two accumulators alternate
            f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
            for (int t = 0; t < c.S / 4; t += 8) {  // S / 4 is a multiple of 8
                float2 ev[8], uv[8];
                float xv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int n = 4 * (t + u) + kk;
                    ev[u] = E[idx];
                    uv[u] = Fa[__builtin_bitreverse32(n) >> (32 - bits)];
                    xv[u] = col < kMomQ ? X[n * kMomQ + col] : 0.0f;
                    idx = (idx + step) & (c.N - 1);
                }
#pragma unroll
                for (int u = 0; u < 8; u += 2) {
                    const float2 e0 = cmulc(ev[u], uv[u]), e1 = cmulc(ev[u + 1], uv[u + 1]);  // e^{-2 pi i m n / N} U[n]
                    acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(part ? e0.y : e0.x, xv[u], acc0, 0, 0, 0);
                    acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(part ? e1.y : e1.x, xv[u + 1], acc1, 0, 0, 0);
                }
            }
            const f32x4 acc = acc0 + acc1;
            // row 4 kk + r of C: band row j = 8 T + 2 kk + r / 2, part r % 2; column col = moment q
            if (col < kMomQ) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int jj = 8 * T + 2 * kk + h;
                    if (jj < b.y) {
                        const float2 v = make_float2(acc[2 * h], acc[2 * h + 1]);
                        const int r = col & 3;  // times (-i)^q
                        slot[jj * kMomQ + col] = r == 0 ? v : (r == 1 ? make_float2(v.y, -v.x)
                                                                    : (r == 2 ? make_float2(-v.x, -v.y)
                                                                              : make_float2(-v.y, v.x)));
                    }
                }
            }
        }
    }
}

template <bool FRESNEL>
__global__ __launch_bounds__(256) void vjp_moments_kernel(const Chirp c, const float* __restrict__ pos, int A,
                                                          const float* __restrict__ ctr,
                                                          const float* __restrict__ nrm,
                                                          const float2* __restrict__ w, int S,
                                                          const float2* __restrict__ slots,
                                                          const int2* __restrict__ bands, int per_range,
                                                          float* __restrict__ partial) {
    __shared__ uint4 stage[kMomGroup * kMomSlotU4];
    __shared__ int2 bd[kMomGroup];
    const int tid = threadIdx.x;
    const int lin = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int tile = lin % gridDim.x, range = lin / gridDim.x;
    const int s = tile * kMomSurfels + tid;
    const bool s_ok = s < S;
    const int sc = min(s, S - 1);
    const int a_begin = range * per_range, a_end = min(A, a_begin + per_range);
    const float cx = ctr[3 * sc], cy = ctr[3 * sc + 1], cz = ctr[3 * sc + 2];
    const float nx = nrm[3 * sc], ny = nrm[3 * sc + 1], nz = nrm[3 * sc + 2];
    const float2 ws = w[sc];
    const float n0 = 0.5f * (float)(c.S - 1), hh = 0.5f * (float)c.S;
    const float t_per_delta = c.pi_N * (float)c.S;  // t = pi delta S / N
    const float rho_per_delta = -c.two_pi_N * n0;   // arg rho = -2 pi delta n0 / N
    const float inv_N = (float)c.inv_N;
    const float2* T = reinterpret_cast<const float2*>(stage);
    float gacc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};

    for (int g0 = a_begin; g0 < a_end; g0 += kMomGroup) {
        __syncthreads();
        // Apertures past the range stage a clamped aperture's slot (finite values); they are
        // not evaluated.
        stage_to_lds<kMomGroup * kMomSlotU4 / 256>(stage, tid, [&](int e) {
            return reinterpret_cast<const uint4*>(slots) + (size_t)min(g0 + e / kMomSlotU4, A - 1) * kMomSlotU4 +
                   e % kMomSlotU4;
        });
        if (tid < kMomGroup) bd[tid] = bands[min(g0 + tid, A - 1)];
        __syncthreads();
        const int na = min(kMomGroup, a_end - g0);
        // each aperture's position is loaded one aperture ahead, so no step waits for its load
        float px = pos[3 * g0], py = pos[3 * g0 + 1], pz = pos[3 * g0 + 2];
        for (int g = 0; g < na; ++g) {
            const int an = min(g0 + g + 1, A - 1);
            const float qx = pos[3 * an], qy = pos[3 * an + 1], qz = pos[3 * an + 2];
            const Pair q = pair_geometry_t<FRESNEL>(c, px, py, pz, cx, cy, cz, nx, ny, nz);
            const float2* Ta = T + g * kMomSlot;
            const int2 b = bd[g];
            float2 G0, G1;
            if (b.y > 0) {  // uniform over the workgroup
                // in the band by construction; the clamp only keeps a read inside the slot
                const int j = min((q.lobe - b.x) & (c.N - 1), kMomBand - 1);
                const float4* row = reinterpret_cast<const float4*>(Ta + j * kMomQ);
                const float4 m01 = row[0], m23 = row[1], m45 = row[2], m67 = row[3], m89 = row[4];
                const float t = t_per_delta * q.delta;
                float2 P = make_float2(m89.z, m89.w), D = make_float2(0.f, 0.f);
                auto horner = [&](float re, float im) {
                    D = make_float2(fmaf(D.x, t, P.x), fmaf(D.y, t, P.y));
                    P = make_float2(fmaf(P.x, t, re), fmaf(P.y, t, im));
                };
                horner(m89.x, m89.y);
                horner(m67.z, m67.w);
                horner(m67.x, m67.y);
                horner(m45.z, m45.w);
                horner(m45.x, m45.y);
                horner(m23.z, m23.w);
                horner(m23.x, m23.y);
                horner(m01.z, m01.w);
                horner(m01.x, m01.y);
                float sr, cr;
                sincos_small(rho_per_delta * q.delta, &sr, &cr);
                const float2 rho = make_float2(cr, sr);
                G0 = cmul(rho, P);
                const float2 rD = cmul(rho, D);
                G1 = make_float2(fmaf(n0, G0.x, -hh * rD.y), fmaf(n0, G0.y, hh * rD.x));  // n0 G0 + i h rho P'
            } else {
                // wide aperture: the exact sum over the samples, U in the slot
                G0 = G1 = make_float2(0.f, 0.f);
                for (int n = 0; n < c.S; ++n) {
                    const int m = (q.lobe * n) & (c.N - 1);
                    float x = 2.0f * ((float)m + q.delta * (float)n) * inv_N;  // cycles of k0 n / N, times 2
                    x -= 2.0f * rintf(0.5f * x);
                    float sn, cn;
                    sincospi_unit(-x, &sn, &cn);
                    const float2 term = cmul(make_float2(cn, sn), Ta[n]);
                    G0 = G0 + term;
                    G1 = G1 + (float)n * term;
                }
            }
            // g0 = G0 / S, g1 = (2 pi i / (N S)) G1; chained as surfel.slang:294-302
            const float amp = s_ok ? q.amp : 0.0f, amp_cos = s_ok ? q.amp_cos : 0.0f;
            const float2 g0 = c.inv_S * G0;
            const float2 g1 = (c.two_pi_N * c.inv_S) * make_float2(-G1.y, G1.x);
            const float2 z0 = cmulc(q.rot, g0), zz1 = cmulc(q.rot, g1);
            const float2 cw = cmulc(ws, z0);
            const float along = cmulc(ws, zz1).x * c.k0pm_f;
            const float gr = -cw.x * amp * q.inv_r + amp * (cw.y * q.rate - along);
            const float gcos = cw.x * amp_cos;
            gacc[0] += -gr * q.ux + gcos * (q.cosine * q.ux - nx) * q.inv_r;
            gacc[1] += -gr * q.uy + gcos * (q.cosine * q.uy - ny) * q.inv_r;
            gacc[2] += -gr * q.uz + gcos * (q.cosine * q.uz - nz) * q.inv_r;
            gacc[3] += gcos * q.ux;
            gacc[4] += gcos * q.uy;
            gacc[5] += gcos * q.uz;
            gacc[6] += amp * z0.x;
            gacc[7] += amp * z0.y;
            px = qx;
            py = qy;
            pz = qz;
        }
    }
    if (s_ok) {
        float4* dst = reinterpret_cast<float4*>(partial + ((size_t)s * gridDim.y + range) * 8);
        dst[0] = make_float4(gacc[0], gacc[1], gacc[2], gacc[3]);
        dst[1] = make_float4(gacc[4], gacc[5], gacc[6], gacc[7]);
    }
}

hipError_t launch_vjp_moments(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                              const float2* w, int S, const float2* up, void* ws, float* gc, float* gn, float* gw,
                              hipStream_t st) {
    const MomShape v = mom_shape(A, S);
    const MomLayout lay = mom_layout(A, S);
    char* base = static_cast<char*>(ws);
    float* partial = reinterpret_cast<float*>(base + lay.partial);
    float2* slots = reinterpret_cast<float2*>(base + lay.slots);
    int2* bands = reinterpret_cast<int2*>(base + lay.bands);
    const int G = mom_prep_rows(c.N);
    const size_t smem = sizeof(float2) * (c.N + (size_t)G * c.N) + sizeof(float) * kMomQ * (size_t)c.S;
    hipError_t e = allow_lds((const void*)vjp_moments_prep, smem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(vjp_moments_prep, dim3((A + G - 1) / G), dim3(256), smem, st, c, pos, A, ctr, S, up, slots,
                       bands);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (c.fresnel_n > 1.0f)
        hipLaunchKernelGGL((vjp_moments_kernel<true>), dim3(v.tiles, v.ranges), dim3(256), 0, st, c, pos, A, ctr,
                           nrm, w, S, slots, bands, v.per_range, partial);
    else
        hipLaunchKernelGGL((vjp_moments_kernel<false>), dim3(v.tiles, v.ranges), dim3(256), 0, st, c, pos, A, ctr,
                           nrm, w, S, slots, bands, v.per_range, partial);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_vjp_reduce(partial, S, v.ranges, gc, gn, gw, st);
}

}  // namespace dsplat
