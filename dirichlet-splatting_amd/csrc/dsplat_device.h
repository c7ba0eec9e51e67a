// Device math shared by the surfel kernels (gfx950, wave64).
//
// Model (reference: dsplat/kernels/scattering.slang, common.slang, surfel.slang).
// For sensor p, surfel center c and unit normal n:
//   r = |p - c| + 1e-12, u = (p - c) / r, cos = u . n
//   amp = cos^2 F(cos) lambda / (8 pi r), 0 for cos <= 0
//   rot = exp(2 pi i (fc tau + K tau (t0 - tau / 2))), tau = 2 r / c0
//   column[k] = amp rot D(k - k0), k0 = r / spacing (mod N)
// and with S ADC samples, N = S P FFT bins, d = k - k0:
//   D(d)  = nu(d) / S (cot(pi d / N) + i),  nu(d) = sin(pi d / P) exp(-i pi d / P)
//   D'(d) = [(pi / P) q(d) (cot + i) - (pi / N) nu(d) (1 + cot^2)] / S, q = exp(-2 i pi d / P)
//
// Precision plan (every result within 1e-5 relative L2 of float64):
//  * Range, carrier phase and k0 come from float64 per pair: the round-trip phase
//    is ~1.8e3 rad at 120 GHz, so float32 r alone costs 1.5e-4 rad.
//  * k0 = lobe + delta with an integer lobe and |delta| <= 1/2 taken in float64, so
//    every bin offset d = m - delta (m = k - lobe integer) is exact to float32.
//  * cot(pi (m - delta) / N) = (T_m + t) / (1 - T_m t), T_m = cot(pi m / N) from an
//    exactly rounded table and t = tan(pi delta / N): |T_m t| <= 1/2 for m != 0, so
//    the per-bin cotangent is relatively accurate next to the lobe, with one rcp and
//    no angle-addition drift.
//  * nu and q have period P in d; per pair they are P constants from sin and cos of
//    beta_r = pi (r - delta) / P, relatively accurate even for delta -> 0.
//  * The lobe bin m = 0: the forward keeps the exact formula (delta is kept off 0 by
//    1e-8 bins, so nu cot stays finite); derivatives (cancelling O(1/delta) terms)
//    take the lobe bin from lobe_direct() and see a table entry T_0 = 0 in the loop.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dsplat {

constexpr float kPi = 3.14159265358979323846f;
constexpr double kPiD = 3.14159265358979323846;

// Kernel-side chirp constants (by value).
struct Chirp {
    int K, N, S, P;
    int form;            // splat form of the call: 0 sample domain where supported, 1 bin domain
    int Kp;              // K rounded up to a multiple of 16
    double k0pm, fc, slope, t0, c0;
    double two_over_c0;  // 2 / c0: round-trip delay per metre
    double inv_N;        // 1 / N
    double lam_d, fresnel_d;  // wavelength and Fresnel index in float64 (scan synthesis)
    float lam, fresnel_n;
    float inv_S;         // 1 / S
    float pi_P, pi_N;    // pi / P, pi / N
    float two_pi_N;      // 2 pi / N
    float h;             // pi (S - 1) / N
    float k0pm_f;        // 1 / spacing as float
    float rate0, rate1;  // carrier rate = rate0 - rate1 * r (rad / m)
    float c2, c4;        // Taylor coefficients of the lobe ratio
};

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2: linear id b
// runs on XCD b % 8. The returned logical id gives each XCD one contiguous range of logical
// ids (XCD x holds total / 8 of them, one more for x < total % 8), so workgroups that are
// neighbours in logical order (and read the same operands) share one L2.
__device__ __forceinline__ int xcd_remap(int b, int total) {
    constexpr int kXcds = 8;
    const int x = b % kXcds, i = b / kXcds, per = total / kXcds, rem = total % kXcds;
    return x < rem ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}

typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

// FP16 hi + lo of the pair (x, y): hi = RNE in one packed conversion, then lo = RNE(x - hi)
// (the difference is exact in FP32) by v_fma_mix{lo,hi}_f16, which read hi as FP16 from its
// half of the packed register, compute -hi + x in FP32 and round the result into one half of
// lo: three VALU instructions per pair of values. The compiler does not see the asm as a VALU
// write, so it inserts no wait states before an MFMA that reads lo: TO_MFMA adds the two a
// VALU-written VGPR needs before it is an MFMA source (lo stored to LDS needs none).
template <bool TO_MFMA>
__device__ __forceinline__ void split_f16_pair(float x, float y, f16x2& hi, f16x2& lo) {
    hi = f16x2{(_Float16)x, (_Float16)y};
    const uint32_t h = __builtin_bit_cast(uint32_t, hi);
    uint32_t l;
    asm("v_fma_mixlo_f16 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(l) : "v"(h), "v"(x));
    if constexpr (TO_MFMA)
        asm("v_fma_mixhi_f16 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]\n\ts_nop 1"
            : "+v"(l) : "v"(h), "v"(y));
    else
        asm("v_fma_mixhi_f16 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "+v"(l) : "v"(h), "v"(y));
    lo = __builtin_bit_cast(f16x2, l);
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// acc + a b as four FMAs (a separate product and sum would take two more instructions)
__device__ __forceinline__ float2 cmac(float2 acc, float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, fmaf(-a.y, b.y, acc.x)), fmaf(a.x, b.y, fmaf(a.y, b.x, acc.y)));
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // conj(a) * b
    return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float2 operator+(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 operator-(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 operator*(float s, float2 a) { return make_float2(s * a.x, s * a.y); }
__device__ __forceinline__ float2& operator+=(float2& a, float2 b) { a.x += b.x; a.y += b.y; return a; }
__device__ __forceinline__ float2& operator-=(float2& a, float2 b) { a.x -= b.x; a.y -= b.y; return a; }

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// sin(pi x), cos(pi x) for a small reduced argument: Taylor-Horner in y = pi x,
// |y| <= pi / 4 (|x| <= 1/4), truncation < 4e-10.
__device__ __forceinline__ void sincos_small(float y, float* s, float* c) {
    const float y2 = y * y;
    float ps = fmaf(y2, -2.50521084e-8f, 2.75573192e-6f);
    ps = fmaf(y2, ps, -1.98412698e-4f);
    ps = fmaf(y2, ps, 8.33333333e-3f);
    ps = fmaf(y2, ps, -1.66666667e-1f);
    *s = fmaf(y * y2, ps, y);
    float pc = fmaf(y2, 2.08767570e-9f, -2.75573192e-7f);
    pc = fmaf(y2, pc, 2.48015873e-5f);
    pc = fmaf(y2, pc, -1.38888889e-3f);
    pc = fmaf(y2, pc, 4.16666667e-2f);
    pc = fmaf(y2, pc, -0.5f);
    *c = fmaf(y2, pc, 1.0f);
}

// tan(y) for |y| <= 0.025 (y = pi delta / N with N >= 64): y + y^3/3 + 2 y^5/15 + 17 y^7/315.
__device__ __forceinline__ float tan_small(float y) {
    const float y2 = y * y;
    float p = fmaf(y2, 5.39682540e-2f, 1.33333333e-1f);
    p = fmaf(y2, p, 3.33333333e-1f);
    return fmaf(y * y2, p, y);
}

// g(y) = cot(y) - 1/y for |y| <= 0.8 (Laurent series, truncation < 1e-9 relative).
__device__ __forceinline__ float cot_minus_inv(float y) {
    const float y2 = y * y;
    float p = fmaf(y2, -2.19259479e-7f, -2.16440428e-6f);
    p = fmaf(y2, p, -2.13777990e-5f);
    p = fmaf(y2, p, -2.11640212e-4f);
    p = fmaf(y2, p, -2.11640212e-3f);
    p = fmaf(y2, p, -2.22222222e-2f);
    p = fmaf(y2, p, -3.33333333e-1f);
    return y * p;
}

// Geometry of one aperture-surfel pair (surfel.slang:63-85, scattering.slang).
struct Pair {
    float r, inv_r;
    float ux, uy, uz;
    float cosine, amp, amp_cos, rate;
    float2 rot;
    int lobe;     // nearest bin to k0, in [0, N)
    float delta;  // k0 - lobe in [-1/2, 1/2], from float64
};

// sin(pi x), cos(pi x) for |x| <= 1: quarter-turn reduction y = x - q / 2, |y| <= 1/4,
// then the Taylor pair of sincos_small and an exact rotation by q quarter turns.
__device__ __forceinline__ void sincospi_unit(float x, float* s, float* c) {
    const float qf = rintf(2.0f * x);
    const float y = fmaf(-0.5f, qf, x);
    float sy, cy;
    sincos_small(kPi * y, &sy, &cy);
    const int q = (int)qf;
    const bool swap = q & 1;
    float ss = swap ? cy : sy, cc = swap ? sy : cy;
    // quarter turns q = 0, 1, 2, 3 (mod 4): (c, s) -> (c, s), (-s, c), (-c, -s), (s, -c)
    if ((q + 1) & 2) cc = -cc;
    if (q & 2) ss = -ss;
    *s = ss;
    *c = cc;
}

// Branch-free pair geometry: FRESNEL selects the Fresnel factor at compile time, the
// cos <= 0 cut is a select, so callers can interleave it with matrix instructions.
template <bool FRESNEL>
__device__ __forceinline__ Pair pair_geometry_t(const Chirp& c, float px, float py, float pz, float cx, float cy,
                                                float cz, float nx, float ny, float nz) {
    Pair q;
    const double dx = (double)px - (double)cx;
    const double dy = (double)py - (double)cy;
    const double dz = (double)pz - (double)cz;
    // float64 range from a float32 square root and one float64 Newton correction
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const float r0 = __builtin_amdgcn_sqrtf((float)r2);  // v_sqrt_f32 (1 ulp): the Newton step fixes it
    const double r = fma(fma(-(double)r0, (double)r0, r2), (double)(0.5f * rcp(r0)), (double)r0) + 1e-12;
    q.r = (float)r;
    float inv = rcp(q.r);
    inv = fmaf(inv, fmaf(-q.r, inv, 1.0f), inv);
    q.ux = (float)dx * inv;
    q.uy = (float)dy * inv;
    q.uz = (float)dz * inv;
    q.inv_r = inv;
    q.cosine = q.ux * nx + q.uy * ny + q.uz * nz;
    // Scattering amplitude and its cos-derivative (scattering.slang:16-38); 0 for cos <= 0.
    float f = 1.0f, df = 0.0f;
    if constexpr (FRESNEL) {
        const float n = c.fresnel_n;
        const float root = __builtin_amdgcn_sqrtf(n * n - 1.0f + q.cosine * q.cosine);
        const float inv_sum = rcp(root + q.cosine);
        f = (root - q.cosine) * inv_sum;
        df = -2.0f * f * rcp(root);
    }
    const float path = c.lam * q.inv_r * (1.0f / (8.0f * kPi));
    const bool lit = q.cosine > 0.0f;
    q.amp = lit ? path * q.cosine * q.cosine * f : 0.0f;
    q.amp_cos = lit ? path * (2.0f * q.cosine * f + q.cosine * q.cosine * df) : 0.0f;
    // Chirp phase in cycles, reduced in float64 (scattering.slang:40-45).
    const double tau = r * c.two_over_c0;
    double cyc = c.fc * tau + c.slope * tau * (c.t0 - 0.5 * tau);
    cyc -= rint(cyc);
    float sp, cp;
    sincospi_unit(2.0f * (float)cyc, &sp, &cp);
    q.rot = make_float2(cp, sp);
    q.rate = fmaf(-c.rate1, q.r, c.rate0);
    // Peak bin k0 = r / spacing wrapped to [0, N), split as lobe + delta. A product by 1/N
    // that lands one ulp off an integer leaves k0 just below 0 or at N: rint and the
    // wrap below keep lobe in [0, N) and |delta| <= 1/2 either way.
    double k0 = r * c.k0pm;
    const double nf = (double)c.N;
    k0 -= nf * floor(k0 * c.inv_N);
    const double lobe = rint(k0);
    q.delta = (float)(k0 - lobe);
    int L = (int)lobe;
    q.lobe = L >= c.N ? L - c.N : (L < 0 ? L + c.N : L);
    return q;
}

__device__ __forceinline__ Pair pair_geometry(const Chirp& c, float px, float py, float pz,
                                              float cx, float cy, float cz,
                                              float nx, float ny, float nz) {
    return c.fresnel_n > 1.0f ? pair_geometry_t<true>(c, px, py, pz, cx, cy, cz, nx, ny, nz)
                              : pair_geometry_t<false>(c, px, py, pz, cx, cy, cz, nx, ny, nz);
}

// sin and cos of beta = pi (rho - delta) / P for every bin class r = k mod P, where
// rho = (r - lobe) mod P is the class of the offset m = k - lobe that nu and q depend
// on (both have period P in d = m - delta). Relative-accurate: the small rotation by
// pi delta / P is applied to exact constants sin, cos (pi rho / P), P in {1, 2, 4, 8}.
template <int P>
__device__ __forceinline__ void class_angles(float delta, int lobe, float* sb, float* cb) {
    float s0, c0;
    if (P == 1) {
        sincospif(delta, &s0, &c0);  // |pi delta| <= pi / 2: the library routine
    } else {
        sincos_small(kPi * delta / (float)P, &s0, &c0);  // |pi delta / P| <= pi / 4
    }
#pragma unroll
    for (int r = 0; r < P; ++r) {
        const int rho = (r - lobe) & (P - 1);
        float sa = 0.0f, ca = 1.0f;
        if (P == 2) {
            sa = rho ? 1.0f : 0.0f;
            ca = rho ? 0.0f : 1.0f;
        } else if (P == 4) {
            const float ts[4] = {0.f, 0.70710678f, 1.f, 0.70710678f};
            const float tc[4] = {1.f, 0.70710678f, 0.f, -0.70710678f};
#pragma unroll
            for (int i = 1; i < 4; ++i)
                if (rho == i) { sa = ts[i]; ca = tc[i]; }
        } else if (P == 8) {
            const float ts[8] = {0.f, 0.38268343f, 0.70710678f, 0.92387953f, 1.f, 0.92387953f, 0.70710678f, 0.38268343f};
            const float tc[8] = {1.f, 0.92387953f, 0.70710678f, 0.38268343f, 0.f, -0.38268343f, -0.70710678f, -0.92387953f};
#pragma unroll
            for (int i = 1; i < 8; ++i)
                if (rho == i) { sa = ts[i]; ca = tc[i]; }
        }
        sb[r] = sa * c0 - ca * s0;
        cb[r] = ca * c0 + sa * s0;
    }
}

// D(-delta) and D'(-delta) at the lobe bin, without cancellation:
// D = R e^{-i h x}, D' = e^{-i h x} (R' - i h R), x = -delta,
// R = sin(a x) / (S sin(b x)), R' / R = a g(a x) - b g(b x), g(y) = cot y - 1 / y.
// (common.slang:10-33 evaluates R' by a quotient that cancels for small x.)
__device__ __forceinline__ void lobe_direct(const Chirp& c, float delta, float2* D, float2* Dp) {
    const float x = -delta;
    const float a = c.pi_P, b = c.pi_N;
    float R;
    if (fabsf(x) < 1e-3f) {
        const float x2 = x * x;
        R = fmaf(x2, fmaf(x2, c.c4, c.c2), 1.0f);
    } else {
        R = sinf(a * x) / ((float)c.S * sinf(b * x));
    }
    float ga;
    const float ax = a * x;
    if (fabsf(ax) <= 0.8f) ga = cot_minus_inv(ax);
    else ga = cosf(ax) / sinf(ax) - 1.0f / ax;
    const float gb = cot_minus_inv(b * x);
    const float Rp = R * (a * ga - b * gb);
    float sh, ch;
    sincosf(c.h * x, &sh, &ch);
    const float2 ph = make_float2(ch, -sh);
    *D = R * ph;
    *Dp = cmul(ph, make_float2(Rp, -c.h * R));
}

// Exactly rounded cot(pi m / N) table entry for index i = m + N, i in [0, 2N).
// m = 0 holds `poison`: 1e30 for the forward (then cot -> -1/t, the exact limit),
// 0 for derivative kernels (cot -> t, removed again after the bin loop).
__device__ __forceinline__ float cot_entry(int i, int N, float poison) {
    const int m = i - N;
    if (m == 0) return poison;
    if (m == -N) return 0.0f;  // never indexed
    const double ang = kPiD * (double)m / (double)N;
    return (float)(cos(ang) / sin(ang));
}

// Per-bin cotangent of pi (m - delta) / N from table entry T_m and t = tan(pi delta / N).
__device__ __forceinline__ float cot_bin(float Tm, float t) {
    return (Tm + t) * rcp(fmaf(-Tm, t, 1.0f));
}

// D(m - delta) and D'(m - delta) at one integer bin offset m (any sign), relatively
// accurate: the lobe (m = 0 mod N) directly, every other bin in cot form with
// T_m = cot(pi m / N) evaluated here rather than read from a table.
__device__ __forceinline__ void kernel_at(const Chirp& c, int m, float delta, float2* D, float2* Dp) {
    int mm = m % c.N;
    if (mm < 0) mm += c.N;
    if (mm == 0) {
        lobe_direct(c, delta, D, Dp);
        return;
    }
    if (mm >= c.N / 2 + (c.N & 1)) mm -= c.N;  // into [-N/2, N/2)
    float sm, cm;
    sincospif((float)mm / (float)c.N, &sm, &cm);
    const float cot = cot_bin(cm / sm, tan_small(c.pi_N * delta));
    int r = mm % c.P;
    if (r < 0) r += c.P;
    float sb, cb;
    sincospif(((float)r - delta) / (float)c.P, &sb, &cb);
    const float2 nu = make_float2(sb * cb, -sb * sb);
    const float2 q = make_float2(fmaf(-2.0f * sb, sb, 1.0f), -2.0f * sb * cb);
    *D = c.inv_S * cmul(nu, make_float2(cot, 1.0f));
    *Dp = c.inv_S * (c.pi_P * cmul(q, make_float2(cot, 1.0f)) - (c.pi_N * fmaf(cot, cot, 1.0f)) * nu);
}

// d amp / d c_j = amp u_j / r + amp_cos (cos u_j - n_j) / r (correlate.slang:223).
__device__ __forceinline__ void amp_gradient(const Pair& q, float nx, float ny, float nz, float* da) {
    da[0] = (q.amp * q.ux + q.amp_cos * (q.cosine * q.ux - nx)) * q.inv_r;
    da[1] = (q.amp * q.uy + q.amp_cos * (q.cosine * q.uy - ny)) * q.inv_r;
    da[2] = (q.amp * q.uz + q.amp_cos * (q.cosine * q.uz - nz)) * q.inv_r;
}

// ---- sample-domain helpers (n = 16 j + i power factorizations) ----

__device__ __forceinline__ float2 csq(float2 a) { return make_float2(a.x * a.x - a.y * a.y, 2.0f * a.x * a.y); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// v_permlane32_swap: lanes 32-63 of a trade places with lanes 0-31 of b. With a = x and
// b = y held by both lane halves, afterwards a = (x of half 0, y of half 0) and
// b = (x of half 1, y of half 1) across the two halves.
__device__ __forceinline__ void swap_halves(float& a, float& b) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}

// e^{2 pi i m / N} for m in [0, N), exactly rounded from float64.
__device__ __forceinline__ void fill_twiddles(float2* E, int N, int tid, int nthreads) {
    for (int m = tid; m < N; m += nthreads) {
        double s, c;
        sincospi(2.0 * (double)m / (double)N, &s, &c);
        E[m] = make_float2((float)c, (float)s);
    }
}

// e^{2 pi i (q lobe + q delta) / N} = E[q lobe mod N] * e^{2 pi i q delta / N}, |q delta| <= q / 2.
// The small rotation needs |2 pi q delta / N| <= pi / 4: q <= N / 4.
__device__ __forceinline__ float2 rotation(const Chirp& ch, const float2* E, int lobe, float delta, int q) {
    float s, c;
    sincos_small(ch.two_pi_N * (float)q * delta, &s, &c);
    return cmul(E[(q * lobe) & (ch.N - 1)], make_float2(c, s));
}

}  // namespace dsplat
