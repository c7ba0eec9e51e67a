// Surfel splat: forward profiles and VJP (reference: dsplat/kernels/surfel.slang:87-311,
// bindings dsplat/kernels/surfel.py:58-137). See dsplat_device.h for the math.
#include "dsplat_device.h"
#include "dsplat_internal.h"

namespace dsplat {

// ---------------------------------------------------------------------------------------
// Forward. A workgroup of 256 threads owns 16 apertures x all bins; thread (a, g) owns
// bins k = g + 16 j, j < KB, all of one bin class g mod P (P divides 16). Surfels are
// streamed in chunks of SC: every pair's float64 prologue is computed once into LDS,
// then each thread accumulates its bins from a broadcast read of the pair:
//   out[a,k] = sum_s g_{s,r} (cot_{s,k} + i),  g_{s,r} = w_s amp rot nu_r / S.
// Per (a, k, s): one LDS table read, 1 add, 1 fma, 1 rcp, 1 mul, 2 fma.
// ---------------------------------------------------------------------------------------
constexpr int kFwdApertures = 16;

template <int P>
struct FwdShape {
    static constexpr int SC = 128 / P;  // surfels per chunk: LDS pair data stays 16 KB
};

template <int P, int KB>
__global__ __launch_bounds__(256) void render_kernel(const Chirp c, const float* __restrict__ pos, int A,
                                                     const float* __restrict__ ctr,
                                                     const float* __restrict__ nrm,
                                                     const float2* __restrict__ w, int S,
                                                     float2* __restrict__ out) {
    constexpr int AT = kFwdApertures;
    constexpr int SC = FwdShape<P>::SC;
    extern __shared__ float smem[];
    float* T = smem;                                   // [2N]
    int* p_off = reinterpret_cast<int*>(T + 2 * c.N);  // [AT * SC] table base N - lobe
    float* p_t = reinterpret_cast<float*>(p_off + AT * SC);  // [AT * SC] tan(pi delta / N)
    float2* p_g = reinterpret_cast<float2*>(p_t + AT * SC);  // [AT * SC * P]

    const int tid = threadIdx.x;
    const int al = tid >> 4, g = tid & 15, cls = g % P;
    const int a0 = blockIdx.x * AT;
    for (int i = tid; i < 2 * c.N; i += 256) T[i] = cot_entry(i, c.N, 1e30f);

    float2 acc[KB];
#pragma unroll
    for (int j = 0; j < KB; ++j) acc[j] = make_float2(0.f, 0.f);
    float2 cc = make_float2(0.f, 0.f);

    for (int s0 = 0; s0 < S; s0 += SC) {
        __syncthreads();
        for (int qi = tid; qi < AT * SC; qi += 256) {
            const int ap = qi / SC, sp = qi - ap * SC;
            const int a = min(a0 + ap, A - 1), s = s0 + sp;
            int off = c.N;
            float t = 0.5f;
            float2 gr[P];
#pragma unroll
            for (int r = 0; r < P; ++r) gr[r] = make_float2(0.f, 0.f);
            if (s < S) {
                const Pair q = pair_geometry(c, pos[3 * a], pos[3 * a + 1], pos[3 * a + 2], ctr[3 * s],
                                             ctr[3 * s + 1], ctr[3 * s + 2], nrm[3 * s], nrm[3 * s + 1],
                                             nrm[3 * s + 2]);
                if (q.amp != 0.0f) {
                    // keep delta off 0 so nu_0 cot_0 stays a finite product (error <= 2e-8)
                    const float dl = fabsf(q.delta) < 1e-8f ? 1e-8f : q.delta;
                    off = c.N - q.lobe;
                    t = tan_small(c.pi_N * dl);
                    float sb[P], cb[P];
                    class_angles<P>(dl, q.lobe, sb, cb);
                    const float2 cw = (q.amp * c.inv_S) * cmul(w[s], q.rot);
#pragma unroll
                    for (int r = 0; r < P; ++r) gr[r] = cmul(cw, make_float2(sb[r] * cb[r], -sb[r] * sb[r]));
                }
            }
            p_off[qi] = off;
            p_t[qi] = t;
#pragma unroll
            for (int r = 0; r < P; ++r) p_g[qi * P + r] = gr[r];
        }
        __syncthreads();
        const int ns = min(SC, S - s0);
        const int row = al * SC;
        for (int sp = 0; sp < ns; ++sp) {
            const int base = p_off[row + sp] + g;
            const float t = p_t[row + sp];
            const float2 gr = p_g[(row + sp) * P + cls];
            cc += gr;
#pragma unroll
            for (int j = 0; j < KB; ++j) {
                const float cot = cot_bin(T[base + 16 * j], t);
                acc[j].x = fmaf(gr.x, cot, acc[j].x);
                acc[j].y = fmaf(gr.y, cot, acc[j].y);
            }
        }
    }
    const int a = a0 + al;
    if (a < A) {
#pragma unroll
        for (int j = 0; j < KB; ++j) {
            const int k = g + 16 * j;
            if (k < c.K) out[(size_t)a * c.K + k] = make_float2(acc[j].x - cc.y, acc[j].y + cc.x);
        }
    }
}

// ---------------------------------------------------------------------------------------
// JVP (surfel.slang:146-221): the forward's layout with per-class coefficients
//   term = h1_r (cot + i) + h2_r (1 + cot^2),  h1_r = (e1 nu_r + (pi/P) e2 q_r) / S,
//   h2_r = -(pi/N) e2 nu_r / S,   e1 = rot ((damp + i amp dphase) w + amp dw),
//   e2 = -amp k0pm dr rot w.
// The derivative table has T_0 = 0, so the lobe bin first receives
// h1 (t + i) + h2 (1 + t^2); the thread owning that bin replaces it by the direct
// e1 D(-delta) + e2 D'(-delta), in pair order in its own LDS slots: deterministic.
// Per (a, k, s): one LDS table read, add, fma, rcp, mul, mul, 4 fma.
// ---------------------------------------------------------------------------------------
template <int P>
struct JvpShape {
    static constexpr int SC = 64 / P;
};

template <int P, int KB>
__global__ __launch_bounds__(256) void jvp_kernel(const Chirp c, const float* __restrict__ pos, int A,
                                                  const float* __restrict__ ctr,
                                                  const float* __restrict__ nrm,
                                                  const float2* __restrict__ w, const float* __restrict__ dctr,
                                                  const float* __restrict__ dnrm, const float2* __restrict__ dw,
                                                  int S, float2* __restrict__ out) {
    constexpr int AT = kFwdApertures;
    constexpr int SC = JvpShape<P>::SC;
    extern __shared__ float smem[];
    float* T = smem;                                                // [2N]
    float2* corr = reinterpret_cast<float2*>(T + 2 * c.N);          // [AT][Kp] lobe corrections
    float2* p_h = corr + AT * c.Kp;                                 // [AT * SC][2P]: h1_r, h2_r
    float2* p_fix = p_h + AT * SC * 2 * P;                          // [AT * SC] lobe correction
    int* p_off = reinterpret_cast<int*>(p_fix + AT * SC);           // [AT * SC]
    int* p_lobe = p_off + AT * SC;                                  // [AT * SC] lobe, -1 if unobserved
    float* p_t = reinterpret_cast<float*>(p_lobe + AT * SC);        // [AT * SC]

    const int tid = threadIdx.x;
    const int al = tid >> 4, g = tid & 15, cls = g % P;
    const int a0 = blockIdx.x * AT;
    for (int i = tid; i < 2 * c.N; i += 256) T[i] = cot_entry(i, c.N, 0.0f);
    for (int i = tid; i < AT * c.Kp; i += 256) corr[i] = make_float2(0.f, 0.f);

    float2 acc[KB];
#pragma unroll
    for (int j = 0; j < KB; ++j) acc[j] = make_float2(0.f, 0.f);
    float2 cc = make_float2(0.f, 0.f);

    for (int s0 = 0; s0 < S; s0 += SC) {
        __syncthreads();
        for (int qi = tid; qi < AT * SC; qi += 256) {
            const int ap = qi / SC, sp = qi - ap * SC;
            const int a = min(a0 + ap, A - 1), s = s0 + sp;
            int off = c.N, lobe = -1;
            float t = 0.5f;
            float2 h1[P], h2[P], fix = make_float2(0.f, 0.f);
#pragma unroll
            for (int r = 0; r < P; ++r) { h1[r] = make_float2(0.f, 0.f); h2[r] = make_float2(0.f, 0.f); }
            if (s < S) {
                const float nx = nrm[3 * s], ny = nrm[3 * s + 1], nz = nrm[3 * s + 2];
                const Pair q = pair_geometry(c, pos[3 * a], pos[3 * a + 1], pos[3 * a + 2], ctr[3 * s],
                                             ctr[3 * s + 1], ctr[3 * s + 2], nx, ny, nz);
                if (q.amp != 0.0f) {
                    const float dcx = dctr[3 * s], dcy = dctr[3 * s + 1], dcz = dctr[3 * s + 2];
                    const float dr = -(q.ux * dcx + q.uy * dcy + q.uz * dcz);
                    const float dcos = ((q.cosine * q.ux - nx) * dcx + (q.cosine * q.uy - ny) * dcy +
                                        (q.cosine * q.uz - nz) * dcz) * q.inv_r +
                                       q.ux * dnrm[3 * s] + q.uy * dnrm[3 * s + 1] + q.uz * dnrm[3 * s + 2];
                    const float damp = -q.amp * dr * q.inv_r + q.amp_cos * dcos;
                    const float2 ws = w[s];
                    const float2 e1 = cmul(q.rot, cmul(make_float2(damp, q.amp * q.rate * dr), ws) + q.amp * dw[s]);
                    const float2 e2 = (-q.amp * c.k0pm_f * dr) * cmul(q.rot, ws);
                    off = c.N - q.lobe;
                    t = tan_small(c.pi_N * q.delta);
                    float sb[P], cb[P];
                    class_angles<P>(q.delta, q.lobe, sb, cb);
#pragma unroll
                    for (int r = 0; r < P; ++r) {
                        const float2 nu = make_float2(sb[r] * cb[r], -sb[r] * sb[r]);
                        const float2 qr = make_float2(fmaf(-2.0f * sb[r], sb[r], 1.0f), -2.0f * sb[r] * cb[r]);
                        h1[r] = c.inv_S * (cmul(e1, nu) + c.pi_P * cmul(e2, qr));
                        h2[r] = (-c.pi_N * c.inv_S) * cmul(e2, nu);
                    }
                    if (q.lobe < c.K) {
                        lobe = q.lobe;
                        float2 D, Dp;
                        lobe_direct(c, q.delta, &D, &Dp);
                        const int rl = q.lobe % P;
                        float2 h1l = h1[0], h2l = h2[0];
#pragma unroll
                        for (int r = 1; r < P; ++r)
                            if (r == rl) { h1l = h1[r]; h2l = h2[r]; }
                        const float2 seen = cmul(h1l, make_float2(t, 1.0f)) + fmaf(t, t, 1.0f) * h2l;
                        fix = cmul(e1, D) + cmul(e2, Dp) - seen;
                    }
                }
            }
            p_off[qi] = off;
            p_lobe[qi] = lobe;
            p_t[qi] = t;
            p_fix[qi] = fix;
#pragma unroll
            for (int r = 0; r < P; ++r) { p_h[(qi * P + r) * 2] = h1[r]; p_h[(qi * P + r) * 2 + 1] = h2[r]; }
        }
        __syncthreads();
        const int ns = min(SC, S - s0);
        const int row = al * SC;
        for (int sp = 0; sp < ns; ++sp) {
            const int qi = row + sp;
            const int base = p_off[qi] + g;
            const float t = p_t[qi];
            const float2 h1 = p_h[(qi * P + cls) * 2], h2 = p_h[(qi * P + cls) * 2 + 1];
            cc += make_float2(h2.x - h1.y, h2.y + h1.x);  // i h1 + h2
            const int lobe = p_lobe[qi];
            if (lobe >= 0 && (lobe & 15) == g) corr[al * c.Kp + lobe] += p_fix[qi];
#pragma unroll
            for (int j = 0; j < KB; ++j) {
                const float cot = cot_bin(T[base + 16 * j], t);
                const float c2 = cot * cot;
                acc[j].x = fmaf(h1.x, cot, fmaf(h2.x, c2, acc[j].x));
                acc[j].y = fmaf(h1.y, cot, fmaf(h2.y, c2, acc[j].y));
            }
        }
    }
    const int a = a0 + al;
    if (a < A) {
#pragma unroll
        for (int j = 0; j < KB; ++j) {
            const int k = g + 16 * j;
            if (k < c.K) out[(size_t)a * c.K + k] = acc[j] + cc + corr[al * c.Kp + k];
        }
    }
}

// ---------------------------------------------------------------------------------------
// VJP. A workgroup owns 32 apertures (upstream tile [Kp][32] in LDS) and a chunk of
// surfels; each half-wave takes one surfel, lane = aperture, and reduces all bins of
// its pair in registers as P class sums (surfel.slang:260-282 in cot form):
//   F_r = sum_{k in r} cot_k u_k,  G_r = sum_{k in r} cot_k^2 u_k,  U_r = sum_{k in r} u_k
// g0 = sum conj(D) u, g1 = sum conj(D') u follow per class; the lobe bin enters
// directly. The 8 surfel gradients are chained per pair, summed over the half-wave and
// written as one partial per (surfel, aperture block); vjp_reduce sums those in float64.
// Per (a, k, s): two LDS reads, 1 add, 3 fma... (add, fma, rcp, mul, 2 fma, mul, 2 fma).
// ---------------------------------------------------------------------------------------
constexpr int kVjpApertures = 32;

template <int P>
__global__ __launch_bounds__(256) void vjp_kernel(const Chirp c, const float* __restrict__ pos, int A,
                                                  const float* __restrict__ ctr,
                                                  const float* __restrict__ nrm,
                                                  const float2* __restrict__ w, int S,
                                                  const float2* __restrict__ up, int chunk,
                                                  float* __restrict__ partial) {
    constexpr int AT = kVjpApertures;
    extern __shared__ float smem[];
    float* T = smem;                                     // [2N]
    float2* U = reinterpret_cast<float2*>(T + 2 * c.N);  // [Kp][AT]
    float2* Us = U + c.Kp * AT;                          // [P][AT]

    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63, half = lane >> 5, al = lane & 31;
    const int blk = blockIdx.x, nblk = gridDim.x;
    const int a0 = blk * AT;
    for (int i = tid; i < 2 * c.N; i += 256) T[i] = cot_entry(i, c.N, 0.0f);
    for (int i = tid; i < AT * c.Kp; i += 256) {
        const int ap = i / c.Kp, k = i - ap * c.Kp;
        const int a = a0 + ap;
        U[k * AT + ap] = (a < A && k < c.K) ? up[(size_t)a * c.K + k] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    if (tid < P * AT) {
        const int r = tid / AT, ap = tid - r * AT;
        float2 s = make_float2(0.f, 0.f);
        for (int k = r; k < c.Kp; k += P) s += U[k * AT + ap];
        Us[r * AT + ap] = s;
    }
    __syncthreads();

    const int a = min(a0 + al, A - 1);
    const bool aperture_ok = a0 + al < A;
    const float px = pos[3 * a], py = pos[3 * a + 1], pz = pos[3 * a + 2];
    const int s_begin = blockIdx.y * chunk, s_end = min(S, s_begin + chunk);
    for (int s = s_begin + 2 * wave + half; s < s_end + 0; s += 8) {
        // Every lane of a half-wave has the same s: the loop is uniform per half-wave.
        const float nx = nrm[3 * s], ny = nrm[3 * s + 1], nz = nrm[3 * s + 2];
        const Pair q = pair_geometry(c, px, py, pz, ctr[3 * s], ctr[3 * s + 1], ctr[3 * s + 2], nx, ny, nz);
        const float amp = aperture_ok ? q.amp : 0.0f;
        const float t = tan_small(c.pi_N * q.delta);
        const int tb = c.N - q.lobe;
        float2 F[P], G[P];
#pragma unroll
        for (int r = 0; r < P; ++r) { F[r] = make_float2(0.f, 0.f); G[r] = make_float2(0.f, 0.f); }
        for (int k0 = 0; k0 < c.Kp; k0 += 16) {
            const float* Tk = T + tb + k0;
            const float2* Uk = U + k0 * AT + al;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const float cot = cot_bin(Tk[j], t);
                const float2 u = Uk[j * AT];
                F[j % P].x = fmaf(cot, u.x, F[j % P].x);
                F[j % P].y = fmaf(cot, u.y, F[j % P].y);
                const float c2 = cot * cot;
                G[j % P].x = fmaf(c2, u.x, G[j % P].x);
                G[j % P].y = fmaf(c2, u.y, G[j % P].y);
            }
        }
        // The lobe bin saw cot = t (table entry 0): take it out, add the direct kernel.
        const int rl = q.lobe % P;
        const float2 ul = q.lobe < c.K ? U[q.lobe * AT + al] : make_float2(0.f, 0.f);
        float2 D, Dp;
        lobe_direct(c, q.delta, &D, &Dp);
        float2 g0 = cmulc(D, ul), g1 = cmulc(Dp, ul);
        float sb[P], cb[P];
        class_angles<P>(q.delta, q.lobe, sb, cb);
        const float t2 = t * t;
#pragma unroll
        for (int r = 0; r < P; ++r) {
            float2 Ur = Us[r * AT + al];
            float2 Fr = F[r], Gr = G[r];
            if (r == rl) {
                Ur -= ul;
                Fr.x = fmaf(-t, ul.x, Fr.x);
                Fr.y = fmaf(-t, ul.y, Fr.y);
                Gr.x = fmaf(-t2, ul.x, Gr.x);
                Gr.y = fmaf(-t2, ul.y, Gr.y);
            }
            const float2 nu = make_float2(sb[r] * cb[r], -sb[r] * sb[r]);
            const float2 qr = make_float2(fmaf(-2.0f * sb[r], sb[r], 1.0f), -2.0f * sb[r] * cb[r]);
            const float2 Fm = make_float2(Fr.x + Ur.y, Fr.y - Ur.x);  // F - i U
            g0 += c.inv_S * cmulc(nu, Fm);
            g1 += c.inv_S * (c.pi_P * cmulc(qr, Fm) - c.pi_N * cmulc(nu, Ur + Gr));
        }
        // Chain to centers, normals and weights (surfel.slang:294-302).
        const float2 z0 = cmulc(q.rot, g0), z1 = cmulc(q.rot, g1);
        const float2 ws = w[s];
        const float2 cw = cmulc(ws, z0);
        const float along = cmulc(ws, z1).x * c.k0pm_f;
        const float gr = -cw.x * amp * q.inv_r + amp * (cw.y * q.rate - along);
        const float gcos = cw.x * (aperture_ok ? q.amp_cos : 0.0f);
        float v[8];
        v[0] = -gr * q.ux + gcos * (q.cosine * q.ux - nx) * q.inv_r;
        v[1] = -gr * q.uy + gcos * (q.cosine * q.uy - ny) * q.inv_r;
        v[2] = -gr * q.uz + gcos * (q.cosine * q.uz - nz) * q.inv_r;
        v[3] = gcos * q.ux;
        v[4] = gcos * q.uy;
        v[5] = gcos * q.uz;
        v[6] = amp * z0.x;
        v[7] = amp * z0.y;
        // Sum over the 32 apertures of the half-wave (fixed xor tree: deterministic).
#pragma unroll
        for (int o = 16; o >= 1; o >>= 1) {
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] += __shfl_xor(v[j], o, 64);
        }
        if (al == 0) {
            float4* dst = reinterpret_cast<float4*>(partial + ((size_t)s * nblk + blk) * 8);
            dst[0] = make_float4(v[0], v[1], v[2], v[3]);
            dst[1] = make_float4(v[4], v[5], v[6], v[7]);
        }
    }
}

// Sum the partials [S, nblk, 8] of every surfel in float64, one wave per surfel, in a
// fixed order (kernels/surfel.py:136 sums them in double on the host).
__global__ __launch_bounds__(256) void vjp_reduce(const float* __restrict__ partial, int S, int nblk,
                                                  float* __restrict__ gc, float* __restrict__ gn,
                                                  float* __restrict__ gw) {
    const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (s >= S) return;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = lane; b < nblk; b += 64) {
        const float4* src = reinterpret_cast<const float4*>(partial + ((size_t)s * nblk + b) * 8);
        const float4 x = src[0], y = src[1];
        acc[0] += x.x; acc[1] += x.y; acc[2] += x.z; acc[3] += x.w;
        acc[4] += y.x; acc[5] += y.y; acc[6] += y.z; acc[7] += y.w;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] += __shfl_xor(acc[j], o, 64);
    }
    if (lane == 0) {
        gc[3 * s] = (float)acc[0]; gc[3 * s + 1] = (float)acc[1]; gc[3 * s + 2] = (float)acc[2];
        gn[3 * s] = (float)acc[3]; gn[3 * s + 1] = (float)acc[4]; gn[3 * s + 2] = (float)acc[5];
        gw[2 * s] = (float)acc[6]; gw[2 * s + 1] = (float)acc[7];
    }
}

// ---------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------
template <int P>
static hipError_t launch_render_p(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                                  const float2* w, int S, float2* out, hipStream_t st) {
    const int KB = c.Kp / 16;
    const size_t smem = sizeof(float) * 2 * c.N +
                        (size_t)kFwdApertures * FwdShape<P>::SC * (sizeof(int) + sizeof(float) + P * sizeof(float2));
    const dim3 grid((A + kFwdApertures - 1) / kFwdApertures);
#define DSPLAT_RENDER(KBV)                                                                              \
    case KBV:                                                                                           \
        hipLaunchKernelGGL((render_kernel<P, KBV>), grid, dim3(256), smem, st, c, pos, A, ctr, nrm, w, S, out); \
        break;
    switch (KB) {
        DSPLAT_RENDER(1)
        DSPLAT_RENDER(2)
        DSPLAT_RENDER(4)
        DSPLAT_RENDER(8)
        DSPLAT_RENDER(16)
        DSPLAT_RENDER(32)
        default: return hipErrorInvalidValue;
    }
#undef DSPLAT_RENDER
    return hipGetLastError();
}

hipError_t launch_render(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                         const float2* w, int S, float2* out, hipStream_t st) {
    if (use_td(c)) return launch_td_render(c, pos, A, ctr, nrm, w, S, out, st);
    switch (c.P) {
        case 1: return launch_render_p<1>(c, pos, A, ctr, nrm, w, S, out, st);
        case 2: return launch_render_p<2>(c, pos, A, ctr, nrm, w, S, out, st);
        case 4: return launch_render_p<4>(c, pos, A, ctr, nrm, w, S, out, st);
        case 8: return launch_render_p<8>(c, pos, A, ctr, nrm, w, S, out, st);
        default: return hipErrorInvalidValue;
    }
}

template <int P>
static hipError_t launch_jvp_p(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                               const float2* w, const float* dctr, const float* dnrm, const float2* dw, int S,
                               float2* out, hipStream_t st) {
    constexpr int SC = JvpShape<P>::SC;
    const int KB = c.Kp / 16;
    const size_t smem = sizeof(float) * 2 * c.N + sizeof(float2) * (size_t)kFwdApertures * c.Kp +
                        (size_t)kFwdApertures * SC * (sizeof(float2) * (2 * P + 1) + 2 * sizeof(int) + sizeof(float));
    const dim3 grid((A + kFwdApertures - 1) / kFwdApertures);
#define DSPLAT_JVP(KBV)                                                                                      \
    case KBV: {                                                                                              \
        hipError_t e = allow_lds((const void*)jvp_kernel<P, KBV>, smem);                                     \
        if (e != hipSuccess) return e;                                                                       \
        hipLaunchKernelGGL((jvp_kernel<P, KBV>), grid, dim3(256), smem, st, c, pos, A, ctr, nrm, w, dctr, dnrm, dw, \
                           S, out);                                                                          \
        break;                                                                                               \
    }
    switch (KB) {
        DSPLAT_JVP(1)
        DSPLAT_JVP(2)
        DSPLAT_JVP(4)
        DSPLAT_JVP(8)
        DSPLAT_JVP(16)
        DSPLAT_JVP(32)
        default: return hipErrorInvalidValue;
    }
#undef DSPLAT_JVP
    return hipGetLastError();
}

hipError_t launch_jvp(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm, const float2* w,
                      const float* dctr, const float* dnrm, const float2* dw, int S, float2* out, hipStream_t st) {
    if (use_td(c)) return launch_td_jvp(c, pos, A, ctr, nrm, w, dctr, dnrm, dw, S, out, st);
    switch (c.P) {
        case 1: return launch_jvp_p<1>(c, pos, A, ctr, nrm, w, dctr, dnrm, dw, S, out, st);
        case 2: return launch_jvp_p<2>(c, pos, A, ctr, nrm, w, dctr, dnrm, dw, S, out, st);
        case 4: return launch_jvp_p<4>(c, pos, A, ctr, nrm, w, dctr, dnrm, dw, S, out, st);
        case 8: return launch_jvp_p<8>(c, pos, A, ctr, nrm, w, dctr, dnrm, dw, S, out, st);
        default: return hipErrorInvalidValue;
    }
}

VjpShape vjp_shape(int A, int S) {
    VjpShape v;
    v.nblk = (A + kVjpApertures - 1) / kVjpApertures;
    // Enough workgroups to fill 256 CUs several times over; chunks are multiples of 8.
    int groups = (4096 + v.nblk - 1) / v.nblk;
    const int max_groups = (S + 7) / 8;
    groups = groups < 1 ? 1 : (groups > max_groups ? max_groups : groups);
    int chunk = (S + groups - 1) / groups;
    v.chunk = (chunk + 7) / 8 * 8;
    v.groups = (S + v.chunk - 1) / v.chunk;
    return v;
}

template <int P>
static hipError_t launch_vjp_p(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                               const float2* w, int S, const float2* up, float* partial, float* gc, float* gn,
                               float* gw, hipStream_t st) {
    const VjpShape v = vjp_shape(A, S);
    const size_t smem = sizeof(float) * 2 * c.N + sizeof(float2) * ((size_t)c.Kp * kVjpApertures + P * kVjpApertures);
    hipError_t e = allow_lds((const void*)vjp_kernel<P>, smem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((vjp_kernel<P>), dim3(v.nblk, v.groups), dim3(256), smem, st, c, pos, A, ctr, nrm, w, S, up,
                       v.chunk, partial);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_vjp_reduce(partial, S, v.nblk, gc, gn, gw, st);
}

hipError_t launch_vjp_reduce(const float* partial, int S, int nblk, float* gc, float* gn, float* gw,
                             hipStream_t st) {
    hipLaunchKernelGGL(vjp_reduce, dim3((S + 3) / 4), dim3(256), 0, st, partial, S, nblk, gc, gn, gw);
    return hipGetLastError();
}

hipError_t launch_vjp(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm, const float2* w,
                      int S, const float2* up, float* partial, float* gc, float* gn, float* gw, hipStream_t st) {
    if (use_td(c)) return launch_td_vjp(c, pos, A, ctr, nrm, w, S, up, partial, gc, gn, gw, st);
    switch (c.P) {
        case 1: return launch_vjp_p<1>(c, pos, A, ctr, nrm, w, S, up, partial, gc, gn, gw, st);
        case 2: return launch_vjp_p<2>(c, pos, A, ctr, nrm, w, S, up, partial, gc, gn, gw, st);
        case 4: return launch_vjp_p<4>(c, pos, A, ctr, nrm, w, S, up, partial, gc, gn, gw, st);
        case 8: return launch_vjp_p<8>(c, pos, A, ctr, nrm, w, S, up, partial, gc, gn, gw, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace dsplat
