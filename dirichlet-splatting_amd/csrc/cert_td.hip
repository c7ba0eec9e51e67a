// Certificate correlations, norms and their derivatives by centres and normals in the sample
// domain, for many candidates against one residual (reference: the per-bin loops of
// correlate.slang:174-235, bound at dsplat/kernels/surfel.py:293-315, which the certificate
// ascent of certificate.py:106-201 calls every iteration).
//
// The residual is taken to the sample domain once (U = DFT of its rows) and turned into FP16
// hi + lo operand tiles of four accumulator sets per aperture: the residual sums g0 = sum_n
// w^n U[n] and g1 = sum_n n w^n U[n] (as the VJP), and the norm polynomial f(k0) = 2 Re sum_n
// psi_n w^n with its k0-derivative sum_n n psi_n w^n (as the scan). One workgroup holds 128
// candidates and an aperture range; per aperture pair the VJP's per-pair generation gives the
// B operands, 3 x 2 x (J / 4) FP16 MFMAs the four sums, and the chain rule the 21 values of
// correlate_gradient (without area):
//   corr = sum amp conj(rot) g0,  norm = sum amp^2 f,  and their derivatives through
//   r = |p - c| (d/dc = -u d/dr), cos = u . n (d/dc = (cos u - n) / r, d/dn = u),
//   d amp / dr = -amp / r, d conj(rot) / dr = -i rate conj(rot), d g0 / dk0 = -g1.
#include "dsplat_internal.h"
#include "dsplat_td.h"

namespace dsplat {

constexpr int kCgCandidates = 128;  // per workgroup: 4 waves x 32 lanes
constexpr int kCgPsiShift = 8;      // n psi rows carry 2^(kScanPsiExp - 8) (n < 256)

// Row of a tile of the four-set layout: tiles t < T hold the residual sets (g0, g1) as the
// VJP's rows, tiles t >= T the norm-polynomial sets (f, f'), T = J / 8.
template <int J>
__device__ __forceinline__ void cg_row(int tt, int q, int h, int r, int* mat, int* j) {
    constexpr int T = J / 8;
    int m = 0;
    td_row<J>(tt % T, q, h, r, &m, j);
    *mat = 2 * (tt / T) + m;
}

// Operand tiles [A][2T][hi, lo][64] of one residual: rows of set 0 (U+ + U-, U+ - U- of U and
// n U) scaled by 2^-e_a per aperture; rows of set 1 from psi (times 2^kScanPsiExp) and n psi
// (times 2^(kScanPsiExp - kCgPsiShift)).
template <int J>
__global__ __launch_bounds__(64 * (J / 4)) void cert_operands_kernel(const float2* __restrict__ U, int S, int N,
                                                                     int K, uint4* __restrict__ out,
                                                                     float* __restrict__ scale_exp) {
    constexpr int TT = J / 4;
    __shared__ float red[TT];
    const int a = blockIdx.x, tid = threadIdx.x, tile = tid >> 6, L = tid & 63;
    const int row = L & 31, h = L >> 5;
    int mat, j;
    cg_row<J>(tile, row >> 3, (row >> 2) & 1, row & 3, &mat, &j);
    const int part = row & 1;
    const bool data = mat < 2;
    auto psi = [&](int m) {  // psi_m = conj(phi_m) of the norm polynomial, times 2^kScanPsiExp
        double re = 0.0, im = 0.0;
        if (m == 0) {
            re = (double)K / (2.0 * S);
        } else if (m & 1) {
            double sn, cs;
            sincospi((double)m / (double)N, &sn, &cs);
            const double w = (double)(S - m) / ((double)S * S);
            re = w;
            im = w * cs / sn;
        }
        return make_float2((float)ldexp(re, kScanPsiExp), (float)ldexp(im, kScanPsiExp));
    };
    float v[8];
    float mx = 0.0f;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int Kx = 8 * h + e, i = Kx >> 1, k = Kx & 1;
        const int np = 16 * j + 8 + i, nm = 16 * j + 7 - i;
        float2 up = data ? U[(size_t)a * S + np] : psi(np);
        float2 um = data ? U[(size_t)a * S + nm] : psi(nm);
        if (mat & 1) {  // n-weighted sets; the n psi rows shifted down to stay in FP16 range
            const float fp = data ? (float)np : ldexpf((float)np, -kCgPsiShift);
            const float fm = data ? (float)nm : ldexpf((float)nm, -kCgPsiShift);
            up = fp * up;
            um = fm * um;
        }
        const float2 sum = up + um, dif = up - um;
        const float x = k == 0 ? (part ? sum.y : sum.x) : (part ? dif.x : -dif.y);
        v[e] = x;
        if (data) mx = fmaxf(mx, fabsf(x));
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if (L == 0) red[tile] = mx;
    __syncthreads();
    for (int t = 0; t < TT; ++t) mx = fmaxf(mx, red[t]);  // the residual rows of every tile
    int ex = 0;
    if (mx > 0.0f) frexpf(mx, &ex);
    ex -= 12;
    if (tid == 0) scale_exp[a] = (float)ex;
    const float sc = data ? ldexpf(1.0f, -ex) : 1.0f;
    f16x8 hi, lo;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const float x = v[e] * sc;
        const _Float16 xh = (_Float16)x;
        hi[e] = xh;
        lo[e] = (_Float16)(x - (float)xh);
    }
    uint4* dst = out + (((size_t)a * TT + tile) * 2) * 64 + L;
    dst[0] = __builtin_bit_cast(uint4, hi);
    dst[64] = __builtin_bit_cast(uint4, lo);
}

// grid (candidate tiles of 128, aperture ranges); partial [C][ranges][21]. With live given, a
// tile none of whose candidates is live (live[s] == 0) writes nothing.
template <int J, bool FRESNEL>
__global__ __launch_bounds__(256) void cert_grad_kernel(const Chirp c, const float* __restrict__ pos, int A,
                                                        const float* __restrict__ ctr, const float* __restrict__ nrm,
                                                        int C, const uint4* __restrict__ ops,
                                                        const float* __restrict__ scale_exp, int per_range,
                                                        const int* __restrict__ live, float* __restrict__ partial) {
    constexpr int TT = J / 4;               // tiles per aperture
    constexpr int G = kTdVjpGroup / TT;     // apertures staged per LDS fill (32 KB)
    constexpr int PER = TT * 2 * 64;        // uint4 per aperture
    extern __shared__ float smem[];
    float2* E = reinterpret_cast<float2*>(smem);     // [N]
    float* Sc = reinterpret_cast<float*>(E + c.N);   // [G] scales 2^e_a
    uint4* Aop = reinterpret_cast<uint4*>(Sc + 16);  // [G][TT][2][64]

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, h = lane >> 5, m = lane & 31;
    // the candidate tiles of one aperture range run on one XCD and share its operands in L2
    const int lin = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int tile = lin % gridDim.x, range = lin / gridDim.x;
    const int s = tile * kCgCandidates + 32 * wave + m;
    const bool s_ok = s < C;
    const int sc = min(s, C - 1);
    const int a_begin = range * per_range, a_end = min(A, a_begin + per_range);
    if (live != nullptr && !__syncthreads_or(s_ok && live[sc] != 0)) return;  // uniform
    fill_twiddles(E, c.N, tid, 256);

    const float cx = ctr[3 * sc], cy = ctr[3 * sc + 1], cz = ctr[3 * sc + 2];
    const float nx = nrm[3 * sc], ny = nrm[3 * sc + 1], nz = nrm[3 * sc + 2];
    const float psi_un = ldexpf(1.0f, -kScanPsiExp), dpsi_un = ldexpf(1.0f, kCgPsiShift - kScanPsiExp);
    float acc[kCgValues];
#pragma unroll
    for (int k = 0; k < kCgValues; ++k) acc[k] = 0.0f;

    for (int g0 = a_begin; g0 < a_end; g0 += G) {
        __syncthreads();
        // past the range: a clamped aperture's (finite) operands, zeroed by amp = 0 of those pairs
        stage_to_lds<G * PER / 256>(Aop, tid, [&](int e) {
            return ops + (size_t)min(g0 + e / PER, A - 1) * PER + e % PER;
        });
        if (tid < G) Sc[tid] = ldexpf(1.0f, (int)scale_exp[min(g0 + tid, A - 1)]);
        __syncthreads();
        for (int ga = 0; ga < G; ga += 2) {
            VjpStage<J> cur;
            {
                const int a = g0 + ga + h;  // this lane's pair: (a, s)
                vjp_stage<J, FRESNEL>(c, E, pos, min(a, A - 1), a < a_end && s_ok, cx, cy, cz, nx, ny, nz, cur);
            }
            float2 P[2][4];  // [aperture a / a + 1][set: g0, g1, f, f']
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int q = 0; q < 4; ++q) P[x][q] = make_float2(0.f, 0.f);
#pragma unroll
            for (int tt = 0; tt < TT; ++tt) {
                f32x16 tacc[2];
#pragma unroll
                for (int ap = 0; ap < 2; ++ap) {
#pragma unroll
                    for (int r = 0; r < 16; ++r) tacc[ap][r] = 0.f;
                    const uint4* src = Aop + ((ga + ap) * TT + tt) * 2 * 64 + lane;
                    const f16x8 ah = __builtin_bit_cast(f16x8, src[0]);
                    const f16x8 al = __builtin_bit_cast(f16x8, src[64]);
                    tacc[ap] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, cur.Bh[ap], tacc[ap], 0, 0, 0);
                    tacc[ap] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, cur.Bl[ap], tacc[ap], 0, 0, 0);
                    tacc[ap] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, cur.Bh[ap], tacc[ap], 0, 0, 0);
                }
                // contraction sum_j Z_j M[j]: half h holds the rows of its j set for both apertures
#pragma unroll
                for (int rr = 0; rr < 16; rr += 2) {
                    int mat, j0, j1;
                    cg_row<J>(tt, rr >> 2, 0, rr & 3, &mat, &j0);
                    cg_row<J>(tt, rr >> 2, 1, rr & 3, &mat, &j1);
                    float2 za0 = cur.Z[j0], za1 = cur.Z[j1];
                    swap_halves(za0.x, za1.x);
                    swap_halves(za0.y, za1.y);
                    P[0][mat] = cmac(P[0][mat], za0, make_float2(tacc[0][rr], tacc[0][rr + 1]));
                    P[1][mat] = cmac(P[1][mat], za1, make_float2(tacc[1][rr], tacc[1][rr + 1]));
                }
            }
            // undo the scales, then each half completes the sums of its own pair's aperture
            const float s0 = Sc[ga], s1 = Sc[ga + 1];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float u0 = q < 2 ? s0 : (q == 2 ? psi_un : dpsi_un);
                const float u1 = q < 2 ? s1 : (q == 2 ? psi_un : dpsi_un);
                P[0][q] = u0 * P[0][q];
                P[1][q] = u1 * P[1][q];
                swap_halves(P[0][q].x, P[1][q].x);
                swap_halves(P[0][q].y, P[1][q].y);
            }
            const float2 T0 = P[0][0] + P[1][0], T1 = P[0][1] + P[1][1];
            const float2 T2 = P[0][2] + P[1][2], T3 = P[0][3] + P[1][3];
            const float2 g0v = c.inv_S * T0;
            const float2 g1v = (c.two_pi_N * c.inv_S) * make_float2(-T1.y, T1.x);
            const float f = 2.0f * T2.x;                  // sum over bins of |D|^2
            const float fp = 2.0f * c.two_pi_N * T3.y;    // its derivative by k0
            const float2 z0 = cmulc(cur.rot, g0v), zz1 = cmulc(cur.rot, g1v);
            const float amp = cur.amp, amp2 = amp * amp;
            // dF / dr and dF / dcos of F = amp conj(rot) g0
            const float2 dFdr = make_float2(-amp * cur.inv_r * z0.x + amp * (cur.rate * z0.y - c.k0pm_f * zz1.x),
                                            -amp * cur.inv_r * z0.y + amp * (-cur.rate * z0.x - c.k0pm_f * zz1.y));
            const float2 dFdcos = cur.amp_cos * z0;
            const float Nv = amp2 * f;
            const float dNdr = amp2 * (c.k0pm_f * fp - 2.0f * cur.inv_r * f);
            const float dNdcos = 2.0f * amp * cur.amp_cos * f;
            const float u[3] = {cur.ux, cur.uy, cur.uz};
            const float n[3] = {nx, ny, nz};
            acc[0] += amp * z0.x;
            acc[1] += amp * z0.y;
            acc[2] += Nv;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float tc = (cur.cosine * u[k] - n[k]) * cur.inv_r;  // d cos / d c_k
                acc[3 + k] += -dFdr.x * u[k] + dFdcos.x * tc;
                acc[9 + k] += -dFdr.y * u[k] + dFdcos.y * tc;
                acc[15 + k] += -dNdr * u[k] + dNdcos * tc;
                acc[6 + k] += dFdcos.x * u[k];  // d cos / d n_k = u_k
                acc[12 + k] += dFdcos.y * u[k];
                acc[18 + k] += dNdcos * u[k];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kCgValues; ++k) {
        float x = acc[k], y = acc[k];
        swap_halves(x, y);
        acc[k] = x + y;
    }
    if (h == 0 && s_ok) {
        float* dst = partial + ((size_t)s * gridDim.y + range) * kCgValues;
#pragma unroll
        for (int k = 0; k < kCgValues; ++k) dst[k] = acc[k];
    }
}

// out[s][k] = sum over ranges of partial[s][r][k] in float64 (live candidates only when live
// is given): one workgroup per candidate, thread t summing ranges t, t + 256, ..., then a
// butterfly over each wave and the four waves in order, so the result does not depend on
// timing.
__global__ __launch_bounds__(256) void cert_grad_reduce(const float* __restrict__ partial, int C, int nrange,
                                                        const int* __restrict__ live, float* __restrict__ out) {
    __shared__ double red[4][kCgValues];
    const int s = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (live != nullptr && live[s] == 0) return;  // uniform
    double acc[kCgValues];
#pragma unroll
    for (int k = 0; k < kCgValues; ++k) acc[k] = 0.0;
    const float* src = partial + (size_t)s * nrange * kCgValues;
    for (int r = tid; r < nrange; r += 256) {
#pragma unroll
        for (int k = 0; k < kCgValues; ++k) acc[k] += src[(size_t)r * kCgValues + k];
    }
#pragma unroll
    for (int k = 0; k < kCgValues; ++k) {
        double d = acc[k];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) d += __shfl_xor(d, o, 64);
        if (lane == 0) red[wave][k] = d;
    }
    __syncthreads();
    if (tid < kCgValues)
        out[(size_t)s * kCgValues + tid] = (float)(((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid]);
}

// ---------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------
struct CertOps {
    size_t u, ops, scales, total;  // byte offsets: U [A][S] complex | tiles | exponents
};

static CertOps cert_ops_layout(const Chirp& c, int A) {
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    CertOps o;
    o.u = 0;
    o.ops = up(sizeof(float2) * (size_t)A * c.S);
    o.scales = o.ops + up(sizeof(uint4) * (size_t)A * (c.S / 64) * 2 * 64);  // 2T = S / 64 tiles
    o.total = o.scales + up(sizeof(float) * (size_t)A);
    return o;
}

size_t cert_operands_bytes(const Chirp& c, int A) { return use_td(c) ? cert_ops_layout(c, A).total : 0; }

hipError_t launch_cert_operands(const Chirp& c, const float2* residual, int A, void* operands, hipStream_t st) {
    const CertOps o = cert_ops_layout(c, A);
    char* base = static_cast<char*>(operands);
    float2* U = reinterpret_cast<float2*>(base + o.u);
    uint4* ops = reinterpret_cast<uint4*>(base + o.ops);
    float* scales = reinterpret_cast<float*>(base + o.scales);
    hipError_t e = launch_dft_rows_up(c, residual, A, U, st);
    if (e != hipSuccess) return e;
    if (c.S == 128)
        hipLaunchKernelGGL(cert_operands_kernel<8>, dim3(A), dim3(128), 0, st, U, c.S, c.N, c.K, ops, scales);
    else
        hipLaunchKernelGGL(cert_operands_kernel<16>, dim3(A), dim3(256), 0, st, U, c.S, c.N, c.K, ops, scales);
    return hipGetLastError();
}

struct CertShape {
    int tiles, ranges, per_range;
};

// About 1024 workgroups: aperture ranges of whole staging groups for the candidate tiles.
static CertShape cert_shape(const Chirp& c, int A, int C) {
    const int G = kTdVjpGroup / (c.S / 64);  // apertures per staging group, as the kernel
    CertShape v;
    v.tiles = (C + kCgCandidates - 1) / kCgCandidates;
    const int groups = (A + G - 1) / G;
    int ranges = (1024 + v.tiles - 1) / v.tiles;
    if (ranges > groups) ranges = groups;
    if (ranges < 1) ranges = 1;
    const int per_groups = (groups + ranges - 1) / ranges;
    v.per_range = per_groups * G;
    v.ranges = (A + v.per_range - 1) / v.per_range;
    return v;
}

size_t cert_grad_workspace(const Chirp& c, int A, int C) {
    if (!use_td(c) || A <= 0 || C <= 0) return 0;
    const CertShape v = cert_shape(c, A, C);
    return sizeof(float) * (size_t)kCgValues * C * v.ranges;
}

template <int J>
static hipError_t launch_cert_grad_j(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                                     int C, const void* operands, float* out, void* ws, hipStream_t st,
                                     const int* live) {
    const CertOps o = cert_ops_layout(c, A);
    const char* base = static_cast<const char*>(operands);
    const uint4* ops = reinterpret_cast<const uint4*>(base + o.ops);
    const float* scales = reinterpret_cast<const float*>(base + o.scales);
    float* partial = static_cast<float*>(ws);
    const CertShape v = cert_shape(c, A, C);
    const size_t smem = sizeof(float2) * c.N + sizeof(float) * 16 + sizeof(uint4) * (size_t)kTdVjpGroup * 128;
    const bool fres = c.fresnel_n > 1.0f;
    const void* fn = fres ? (const void*)cert_grad_kernel<J, true> : (const void*)cert_grad_kernel<J, false>;
    hipError_t e = allow_lds(fn, smem);
    if (e != hipSuccess) return e;
    if (fres)
        hipLaunchKernelGGL((cert_grad_kernel<J, true>), dim3(v.tiles, v.ranges), dim3(256), smem, st, c, pos, A, ctr,
                           nrm, C, ops, scales, v.per_range, live, partial);
    else
        hipLaunchKernelGGL((cert_grad_kernel<J, false>), dim3(v.tiles, v.ranges), dim3(256), smem, st, c, pos, A,
                           ctr, nrm, C, ops, scales, v.per_range, live, partial);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(cert_grad_reduce, dim3(C), dim3(256), 0, st, partial, C, v.ranges, live, out);
    return hipGetLastError();
}

hipError_t launch_cert_grad(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm, int C,
                            const void* operands, float* out, void* ws, hipStream_t st, const int* live) {
    if (c.S == 128) return launch_cert_grad_j<8>(c, pos, A, ctr, nrm, C, operands, out, ws, st, live);
    return launch_cert_grad_j<16>(c, pos, A, ctr, nrm, C, operands, out, ws, st, live);
}

}  // namespace dsplat
