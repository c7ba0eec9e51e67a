// Separable windowed Dirichlet atoms on a 1-3-D spectral grid (reference:
// dsplat/kernels/spectral.slang:1-114, bindings dsplat/kernels/spectral.py:1-95).
// All coordinates are bin units; the gradient convention is dc = -dk.
#include "dsplat_device.h"
#include "dsplat_internal.h"

namespace dsplat {

// D(x) and dD/dx of S samples zero-padded to N at an offset x (wrapped here), float32:
// R = sin(a x) / (S sin(b x)) via sinpi, R' = R (a g(a x) - b g(b x)) for |a x| <= 0.8
// (no cancellation), the quotient rule otherwise; D = R e^{-i h x}, D' = e^{-i h x}(R' - i h R).
__device__ __forceinline__ void dirichlet_f32(float delta, int S, int N, float2* D, float2* Dp) {
    const float nf = (float)N;
    const float x = delta - nf * rintf(delta / nf);
    const float a = kPi * (float)S / nf, b = kPi / nf, h = kPi * (float)(S - 1) / nf;
    float R, Rp;
    if (fabsf(x) < 1e-3f) {
        const float c2 = (b * b - a * a) / 6.0f;
        const float c4 = (3.0f * a * a * a * a + 7.0f * b * b * b * b - 10.0f * a * a * b * b) / 360.0f;
        const float x2 = x * x;
        R = fmaf(x2, fmaf(x2, c4, c2), 1.0f);
        Rp = x * fmaf(x2, 4.0f * c4, 2.0f * c2);
    } else {
        float sa, ca, sb, cb;
        sincospif((float)S * x / nf, &sa, &ca);
        sincospif(x / nf, &sb, &cb);
        R = sa / ((float)S * sb);
        if (fabsf(a * x) <= 0.8f) {
            Rp = R * (a * cot_minus_inv(a * x) - b * cot_minus_inv(b * x));
        } else {
            Rp = (a * ca * sb - b * sa * cb) / ((float)S * sb * sb);
        }
    }
    float sh, ch;
    sincosf(h * x, &sh, &ch);
    const float2 ph = make_float2(ch, -sh);
    *D = R * ph;
    *Dp = cmul(ph, make_float2(Rp, -h * R));
}

// One axis of the footprint: rect/Hann/Hamming closed form, Kaiser explicit taps,
// radius truncation (spectral.slang:4-24).
__device__ __forceinline__ void axis_kernel(const SpecGrid& g, const float* __restrict__ taps, int d, float delta,
                                            float2* k, float2* dk) {
    const int N = g.nfft[d], S = g.samples[d];
    const float nf = (float)N;
    delta -= nf * rintf(delta / nf);
    if (g.radius > 0.0f && fabsf(delta) > g.radius) {
        *k = make_float2(0.f, 0.f);
        *dk = make_float2(0.f, 0.f);
        return;
    }
    if (g.window == 3) {  // Kaiser: sum_n t_n exp(-2 pi i n delta / N)
        float2 v = make_float2(0.f, 0.f), dv = make_float2(0.f, 0.f);
        for (int n = 0; n < S; ++n) {
            float s, c;
            const float turns = -2.0f * (float)n * delta / nf;
            sincospif(turns - 2.0f * rintf(0.5f * turns), &s, &c);
            const float2 term = taps[d * g.taps_stride + n] * make_float2(c, s);
            v += term;
            const float omega = -2.0f * kPi * (float)n / nf;
            dv += make_float2(-omega * term.y, omega * term.x);
        }
        *k = v;
        *dk = dv;
        return;
    }
    dirichlet_f32(delta, S, N, k, dk);
    if (g.window == 0) return;
    const float w = g.window == 1 ? 0.5f : 0.46f / 1.08f;
    const float shift = nf / (float)S;
    float2 km, dm, kp, dp;
    dirichlet_f32(delta - shift, S, N, &km, &dm);
    dirichlet_f32(delta + shift, S, N, &kp, &dp);
    *k = *k - w * (km + kp);
    *dk = *dk - w * (dm + dp);
}

// Atom value and -d/dcenter per axis at (atom s, bin m).
__device__ __forceinline__ void footprint(const SpecGrid& g, const float* __restrict__ taps,
                                          const float* __restrict__ ctr, const float* __restrict__ bins, int s, int m,
                                          float2* a, float2* da) {
    float2 k[3], dk[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        k[d] = make_float2(1.f, 0.f);
        dk[d] = make_float2(0.f, 0.f);
    }
    for (int d = 0; d < g.dim; ++d) axis_kernel(g, taps, d, bins[m * g.dim + d] - ctr[s * g.dim + d], &k[d], &dk[d]);
    *a = cmul(cmul(k[0], k[1]), k[2]);
    da[0] = -1.0f * cmul(cmul(dk[0], k[1]), k[2]);
    da[1] = -1.0f * cmul(cmul(k[0], dk[1]), k[2]);
    da[2] = -1.0f * cmul(cmul(k[0], k[1]), dk[2]);
}

// out[m] = sum_s coeff_s a(s, m); with tangents, the JVP sum (spectral.slang:32-47, 94-114).
__global__ __launch_bounds__(128) void spectral_forward_kernel(const SpecGrid g, const float* __restrict__ taps,
                                                               const float* __restrict__ ctr,
                                                               const float2* __restrict__ coeff,
                                                               const float* __restrict__ dctr,
                                                               const float2* __restrict__ dcoeff,
                                                               const float* __restrict__ bins, int S, int M,
                                                               float2* __restrict__ out) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    float2 sum = make_float2(0.f, 0.f);
    for (int s = 0; s < S; ++s) {
        float2 a, da[3];
        footprint(g, taps, ctr, bins, s, m, &a, da);
        if (dctr) {
            float2 t = make_float2(0.f, 0.f);
            for (int d = 0; d < g.dim; ++d) t += dctr[s * g.dim + d] * da[d];
            sum += cmul(t, coeff[s]) + cmul(a, dcoeff[s]);
        } else {
            sum += cmul(a, coeff[s]);
        }
    }
    out[m] = sum;
}

// Dictionary A [M, S] and dA/dcenters [M, S, D] (spectral.slang:76-92).
__global__ __launch_bounds__(256) void spectral_design_kernel(const SpecGrid g, const float* __restrict__ taps,
                                                              const float* __restrict__ ctr,
                                                              const float* __restrict__ bins, int S, int M,
                                                              float2* __restrict__ out, float2* __restrict__ jac) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)M * S) return;
    const int m = (int)(i / S), s = (int)(i % S);
    float2 a, da[3];
    footprint(g, taps, ctr, bins, s, m, &a, da);
    out[i] = a;
    if (jac)
        for (int d = 0; d < g.dim; ++d) jac[i * g.dim + d] = da[d];
}

// Gradients per (atom, span of bins): partial [S][nspan][D + 2] = (gc[D], gb.re, gb.im),
// gc_d = Re(conj(g) b da_d), gb = conj(a) g (spectral.slang:49-74, without atomics).
__global__ __launch_bounds__(64) void spectral_backward_kernel(const SpecGrid g, const float* __restrict__ taps,
                                                               const float* __restrict__ ctr,
                                                               const float2* __restrict__ coeff,
                                                               const float* __restrict__ bins,
                                                               const float2* __restrict__ up, int S, int M, int span,
                                                               float* __restrict__ partial) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const int nspan = gridDim.y, width = g.dim + 2;
    const int start = blockIdx.y * span, stop = min(M, start + span);
    const float2 b = coeff[s];
    float gc[3] = {0.f, 0.f, 0.f};
    float2 gb = make_float2(0.f, 0.f);
    for (int m = start; m < stop; ++m) {
        float2 a, da[3];
        footprint(g, taps, ctr, bins, s, m, &a, da);
        const float2 gm = up[m];
        for (int d = 0; d < g.dim; ++d) gc[d] += cmulc(gm, cmul(b, da[d])).x;
        gb += cmulc(a, gm);
    }
    float* dst = partial + ((size_t)s * nspan + blockIdx.y) * width;
    for (int d = 0; d < g.dim; ++d) dst[d] = gc[d];
    dst[g.dim] = gb.x;
    dst[g.dim + 1] = gb.y;
}

hipError_t launch_spectral_forward(const SpecGrid& g, const float* taps, const float* ctr, const float2* coeff,
                                   const float* dctr, const float2* dcoeff, const float* bins, int S, int M,
                                   float2* out, hipStream_t st) {
    hipLaunchKernelGGL(spectral_forward_kernel, dim3((M + 127) / 128), dim3(128), 0, st, g, taps, ctr, coeff, dctr,
                       dcoeff, bins, S, M, out);
    return hipGetLastError();
}

hipError_t launch_spectral_design(const SpecGrid& g, const float* taps, const float* ctr, const float* bins, int S,
                                  int M, float2* out, float2* jac, hipStream_t st) {
    const size_t blocks = ((size_t)M * S + 255) / 256;
    if (blocks == 0) return hipSuccess;
    if (blocks > 0x7fffffffu) return hipErrorInvalidValue;
    hipLaunchKernelGGL(spectral_design_kernel, dim3((unsigned)blocks), dim3(256), 0, st, g, taps, ctr, bins, S, M,
                       out, jac);
    return hipGetLastError();
}

int spectral_spans(int M) { return (M + kSpectralSpan - 1) / kSpectralSpan; }

hipError_t launch_spectral_backward(const SpecGrid& g, const float* taps, const float* ctr, const float2* coeff,
                                    const float* bins, const float2* up, int S, int M, float* partial, float* out,
                                    hipStream_t st) {
    const int nspan = spectral_spans(M);
    hipLaunchKernelGGL(spectral_backward_kernel, dim3((S + 63) / 64, nspan), dim3(64), 0, st, g, taps, ctr, coeff,
                       bins, up, S, M, kSpectralSpan, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_reduce_partials(partial, S, nspan, g.dim + 2, out, st);
}

}  // namespace dsplat
