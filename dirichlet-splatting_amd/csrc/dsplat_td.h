// Sample-domain pieces shared by the splat VJP, the certificate scan and the certificate
// gradient kernels: the MFMA vector types and group sizes, the direct global-to-LDS staging,
// the accumulator row layout and the per-pair generation of the VJP's B operands.
#pragma once

#include "dsplat_device.h"

namespace dsplat {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

constexpr int kTdApertures = 16;  // forward: apertures per workgroup
constexpr int kTdChunk = 16;      // forward: surfels per LDS chunk
constexpr int kTdVjpSurfels = 128;
constexpr int kTdVjpGroup = 16;   // VJP: apertures staged per LDS fill

constexpr int kScanPsiExp = 10;
constexpr int kScanCandidates = 128;  // per workgroup (4 waves x 32)
constexpr int kScanMinRange = 512;    // apertures per range at least (bounds the partials)

// Straight copy of ITERS x 256 uint4 from global memory to LDS by the direct global -> LDS
// loads of gfx950 (global_load_lds_dwordx4): the data never passes through VGPRs and no load
// waits for another. Element e = tid + 256 it lands in dst[e]: the wave's slot dst + (e & ~63)
// goes in M0 and lane l writes 16 bytes at M0 + 16 l. The caller's barrier follows the wait.
template <int ITERS, typename Src>
__device__ __forceinline__ void stage_to_lds(uint4* dst, int tid, Src src_of) {
    const int wave_base = __builtin_amdgcn_readfirstlane(tid & ~63);  // uniform: kept in SGPRs
#pragma unroll
    for (int it = 0; it < ITERS; ++it) {
        const int e = tid + 256 * it;
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src_of(e),
                                         (__attribute__((address_space(3))) void*)(dst + wave_base + 256 * it), 16,
                                         0, 0);
    }
    __builtin_amdgcn_s_waitcnt(0);
}

// out[r][q] = scale sum_{p < Lin} e^{sign 2 pi i p q / N} X[r][p] for the 16 rows of X (LDS,
// row stride ldx float2) and q < Lout (multiple of 8), one 16 x 8-bin tile per wave at a time
// on v_mfma_f32_16x16x4f32. Lane layouts (gfx950): A[i][k] -> lane i + 16 k, B[k][j] -> lane
// j + 16 k, C[4 (lane / 16) + r][lane % 16] in register r. K index k = 2 dp + pin: sample
// p = 2 t + dp, input part pin; column x = 2 (bin offset) + output part.
// The K loop of one dft16 / dft8_split tile, sum over t < count of a(t) b(t) on
// v_mfma_f32_16x16x4f32 with a(t) = xrow[4 t] and b(t) from E[m + t step]: the LDS reads of 8
// steps are issued before their products and two accumulators alternate, so a step does not
// wait for its own reads and the previous product.
__device__ __forceinline__ f32x4 dft_tile(const float* xrow, const float2* E, int m, int step, int N, float scale,
                                          float sign, int pin, int pout, int count) {
    auto bval = [&](float2 e) {
        const float fr = e.x * scale, fi = sign * e.y * scale;
        return pin ? (pout ? fr : -fi) : (pout ? fi : fr);
    };
    f32x4 d0 = {0.f, 0.f, 0.f, 0.f}, d1 = {0.f, 0.f, 0.f, 0.f};
    int t = 0;
    for (; t + 8 <= count; t += 8) {
        float a[8];
        float2 e[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            a[u] = xrow[4 * (t + u)];
            e[u] = E[m];
            m = (m + step) & (N - 1);
        }
#pragma unroll
        for (int u = 0; u < 8; u += 2) {
            d0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], bval(e[u]), d0, 0, 0, 0);
            d1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u + 1], bval(e[u + 1]), d1, 0, 0, 0);
        }
    }
    for (; t < count; ++t) {
        d0 = __builtin_amdgcn_mfma_f32_16x16x4f32(xrow[4 * t], bval(E[m]), d0, 0, 0, 0);
        m = (m + step) & (N - 1);
    }
    return d0 + d1;
}

__device__ __forceinline__ void dft16(const float2* X, int ldx, int Lin, const float2* E, int N, float sign,
                                      float scale, int Lout, float2* out, int ldo, int rows, int wave,
                                      int lane) {
    const int kk = lane >> 4, x = lane & 15;
    const int pin = kk & 1, dp = kk >> 1, pout = x & 1;
    const float* Xf = reinterpret_cast<const float*>(X);
    for (int ct = wave; ct * 8 < Lout; ct += 4) {
        const int q = ct * 8 + (x >> 1);
        const float* xrow = Xf + 2 * (x * ldx + dp) + pin;
        const f32x4 d = dft_tile(xrow, E, (dp * q) & (N - 1), (2 * q) & (N - 1), N, scale, sign, pin, pout, Lin / 2);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 4 * kk + r;
            if (row < rows) reinterpret_cast<float*>(out + (size_t)row * ldo + q)[pout] = d[r];
        }
    }
}

// ---------------------------------------------------------------------------------------
// VJP
// ---------------------------------------------------------------------------------------
// Tile-local row of the 32 x 32 accumulator held in register rr of lane half h:
// row = 8 (rr / 4) + 4 h + rr % 4 (gfx950 layout of v_mfma_f32_32x32x2f32). Rows hold
// (mat, j, part): for J = 8 one tile, mat = q / 2, j = 4 (q % 2) + 2 h + (r / 2); for J = 16
// tile T = mat, j = 4 q + 2 h + (r / 2); part = r % 2 (q = rr / 4, r = rr % 4).
template <int J>
__device__ __forceinline__ void td_row(int tile, int q, int h, int r, int* mat, int* j) {
    if (J == 8) {
        *mat = q >> 1;
        *j = 4 * (q & 1) + 2 * h + (r >> 1);
    } else {
        *mat = tile;
        *j = 4 * q + 2 * h + (r >> 1);
    }
}

// e^{2 pi i (lobe + delta) / (2 N)} = E[lobe / 2] e^{i pi ((lobe % 2) + delta) / N}: a square
// root of rotation(..., 1), relatively accurate like it (|pi (1 + delta) / N| <= pi / 4).
__device__ __forceinline__ float2 half_rotation(const Chirp& ch, const float2* E, int lobe, float delta) {
    float s, c;
    sincos_small(ch.pi_N * ((float)(lobe & 1) + delta), &s, &c);
    return cmul(E[lobe >> 1], make_float2(c, s));
}

// What one lane carries from the generation of its pair to the contraction and chain rule
// (symmetric split above).
template <int J>
struct VjpStage {
    f16x8 Bh[2], Bl[2];  // [aperture a / a + 1]: FP16 hi / lo of the B operands
    float2 Z[J];         // Y_j = w^(16 j + 7.5) of the own pair
    float2 rot;
    float amp, amp_cos, inv_r, ux, uy, uz, cosine, rate;
};

template <int J, bool FRESNEL>
__device__ __forceinline__ void vjp_stage(const Chirp& c, const float2* E, const float* __restrict__ pos, int a,
                                              bool ok, float cx, float cy, float cz, float nx, float ny, float nz,
                                              VjpStage<J>& st) {
    const Pair q = pair_geometry_t<FRESNEL>(c, pos[3 * a], pos[3 * a + 1], pos[3 * a + 2], cx, cy, cz, nx, ny, nz);
    st.rot = q.rot;
    st.amp = ok ? q.amp : 0.0f;
    st.amp_cos = ok ? q.amp_cos : 0.0f;
    st.inv_r = q.inv_r;
    st.ux = q.ux;
    st.uy = q.uy;
    st.uz = q.uz;
    st.cosine = q.cosine;
    st.rate = q.rate;
    // w = conj(v): W_i = w^(i + 1/2) by two chains over w^2, Y_j = W_7 z1^j by two chains over z1^2
    const float2 w1 = cconj(rotation(c, E, q.lobe, q.delta, 1));
    const float2 z1 = cconj(rotation(c, E, q.lobe, q.delta, 16));
    const float2 w2 = csq(w1);
    float2 W[8];
    W[0] = cconj(half_rotation(c, E, q.lobe, q.delta));
    W[1] = cmul(W[0], w1);
#pragma unroll
    for (int i = 2; i < 8; ++i) W[i] = cmul(W[i - 2], w2);
    const float2 z2 = csq(z1);
    st.Z[0] = W[7];
    st.Z[1] = cmul(W[7], z1);
#pragma unroll
    for (int j = 2; j < J; ++j) st.Z[j] = cmul(st.Z[j - 2], z2);
    // B operands: lane half h supplies K = 8 h + e, i.e. W_i with i = 4 h + e / 2 (Re for even
    // e, Im for odd e). Half h generated aperture a + h; one half swap of (W_q, W_{4 + q})
    // gives each half its i of aperture a and of a + 1.
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
        float xr = W[qq].x, yr = W[4 + qq].x;
        float xi = W[qq].y, yi = W[4 + qq].y;
        swap_halves(xr, yr);
        swap_halves(xi, yi);
        const float vals[2][2] = {{xr, xi}, {yr, yi}};  // [aperture][re, im]
#pragma unroll
        for (int ap = 0; ap < 2; ++ap) {
            f16x2 hi, lo;
            split_f16_pair<true>(vals[ap][0], vals[ap][1], hi, lo);
            st.Bh[ap][2 * qq] = hi.x;
            st.Bh[ap][2 * qq + 1] = hi.y;
            st.Bl[ap][2 * qq] = lo.x;
            st.Bl[ap][2 * qq + 1] = lo.y;
        }
    }
}

}  // namespace dsplat
