// Closed-form bin sums of one aperture-surfel pair, shared by the correlation, scan and
// quasi-Newton search kernels (reference: the per-bin loops of correlate.slang:78-235).
#pragma once

#include "dsplat_device.h"

namespace dsplat {

struct Sums {
    float2 g0, g1;   // sum conj(D) res, sum conj(D') res
    float nrm, rx;   // sum |D|^2, sum Re(conj(D') D)
};

// Bin sums of one pair over the observed bins: T is the derivative cot table [2N] (T_0 = 0),
// R the aperture's residual at bin k in R[k * stride] (zero-padded to Kp), Us its P class
// sums (bins k = r mod P) in Us[r * stride]. GRAD adds g1 and rx.
template <int P, bool GRAD>
__device__ __forceinline__ Sums bin_sums(const Chirp& c, const float* __restrict__ T, const float2* __restrict__ R,
                                         const float2* __restrict__ Us, const Pair& q, int STRIDE) {
    const float t = tan_small(c.pi_N * q.delta);
    const float* Tb = T + (c.N - q.lobe);
    float2 F[P], G[P];
    float Q[P], C1[P], C3[P];
#pragma unroll
    for (int r = 0; r < P; ++r) {
        F[r] = make_float2(0.f, 0.f);
        G[r] = make_float2(0.f, 0.f);
        Q[r] = 0.f;
        C1[r] = 0.f;
        C3[r] = 0.f;
    }
    for (int k0 = 0; k0 < c.K; k0 += 16) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int k = k0 + j;
            if (k < c.K) {  // uniform
                const float cot = cot_bin(Tb[k], t);
                const float2 u = R[k * STRIDE];
                F[j % P].x = fmaf(cot, u.x, F[j % P].x);
                F[j % P].y = fmaf(cot, u.y, F[j % P].y);
                Q[j % P] = fmaf(cot, cot, Q[j % P]);
                if (GRAD) {
                    const float c2 = cot * cot;
                    G[j % P].x = fmaf(c2, u.x, G[j % P].x);
                    G[j % P].y = fmaf(c2, u.y, G[j % P].y);
                    C1[j % P] += cot;
                    C3[j % P] = fmaf(c2, cot, C3[j % P]);
                }
            }
        }
    }
    const bool seen = q.lobe < c.K;
    const int rl = q.lobe % P;
    const float2 ul = seen ? R[q.lobe * STRIDE] : make_float2(0.f, 0.f);
    float2 D, Dp;
    lobe_direct(c, q.delta, &D, &Dp);
    Sums out;
    out.g0 = seen ? cmulc(D, ul) : make_float2(0.f, 0.f);
    out.nrm = seen ? D.x * D.x + D.y * D.y : 0.f;
    out.g1 = seen ? cmulc(Dp, ul) : make_float2(0.f, 0.f);
    out.rx = seen ? cmulc(Dp, D).x : 0.f;
    float sb[P], cb[P];
    class_angles<P>(q.delta, q.lobe, sb, cb);
    const float t2 = t * t, invS2 = c.inv_S * c.inv_S;
#pragma unroll
    for (int r = 0; r < P; ++r) {
        float2 Ur = Us[r * STRIDE];
        float2 Fr = F[r], Gr = G[r];
        float Qr = Q[r], C1r = C1[r], C3r = C3[r];
        float nr = (float)((c.K - r + P - 1) / P);
        if (seen && r == rl) {  // the lobe bin saw cot = t: take it out
            Ur -= ul;
            Fr.x = fmaf(-t, ul.x, Fr.x);
            Fr.y = fmaf(-t, ul.y, Fr.y);
            Qr -= t2;
            nr -= 1.0f;
            if (GRAD) {
                Gr.x = fmaf(-t2, ul.x, Gr.x);
                Gr.y = fmaf(-t2, ul.y, Gr.y);
                C1r -= t;
                C3r = fmaf(-t2, t, C3r);
            }
        }
        const float2 nu = make_float2(sb[r] * cb[r], -sb[r] * sb[r]);
        const float nu2 = sb[r] * sb[r];
        const float2 Fm = make_float2(Fr.x + Ur.y, Fr.y - Ur.x);  // F - i U
        out.g0 += c.inv_S * cmulc(nu, Fm);
        out.nrm += nu2 * invS2 * (nr + Qr);
        if (GRAD) {
            const float2 qr = make_float2(fmaf(-2.0f * sb[r], sb[r], 1.0f), -2.0f * sb[r] * cb[r]);
            out.g1 += c.inv_S * (c.pi_P * cmulc(qr, Fm) - c.pi_N * cmulc(nu, Ur + Gr));
            out.rx += invS2 * (c.pi_P * cmulc(qr, nu).x * (nr + Qr) - c.pi_N * nu2 * (C1r + C3r));
        }
    }
    return out;
}

}  // namespace dsplat
