// FP64 FMCW scan synthesis on the matrix cores: the physics signal chain of
// groundtruth.py:58-90 (float64 beat tones per aperture-surfel echo, summed, zero-padded
// FFT, positive bins, / adc_samples), which never evaluates the Dirichlet kernel.
//
// Per aperture a the ADC record is X[a][n] = sum_s e_{a,s} v_{a,s}^n, n < samples, with the
// echo e = gain e^{2 pi i carrier} b_s and the tone v = e^{2 pi i f}, f = beat / fs = r k0pm / N
// cycles per sample. Splitting n = 16 j + i, v^n = Z_j W_i (Z_j = v^{16 j}, W_i = v^i) turns the
// sum over surfels into one small complex GEMM per aperture: a workgroup of 2 waves owns 8
// apertures; per chunk of 16 surfels each thread writes the record (Z_j, e W_i) of one pair
// to LDS and each wave accumulates its 4 apertures' [2J x 16] real tiles on
// v_mfma_f64_16x16x4f64 as four real products (rows (Z part, j), K = surfel, columns i for the
// real and the imaginary part of e W). The spectrum out[a][k] = (1/samples) sum_n X[a][n]
// e^{-2 pi i n k / N}, k < K, is a float64 DFT from an exactly rounded twiddle table.
#include "dsplat_device.h"
#include "dsplat_internal.h"

namespace dsplat {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kSyApertures = 8;  // apertures per workgroup (2 waves x 4)
constexpr int kSyChunk = 16;     // surfels per chunk: one pair per thread

template <int J>
struct SyRec {
    // doubles per pair record: Z re[J], Z im[J], (e W) re[16], (e W) im[16], plus one so
    // consecutive records start 2 mod 4 dwords apart
    static constexpr int size = 2 * J + 33;
};

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 zsq(double2 a) { return make_double2(fma(a.x, a.x, -a.y * a.y), 2.0 * a.x * a.y); }

// e^{2 pi i x} for x reduced to [-1/2, 1/2]
__device__ __forceinline__ double2 cis_cycles(double x) {
    x -= rint(x);
    double s, c;
    sincospi(2.0 * x, &s, &c);
    return make_double2(c, s);
}

// Per-surfel inputs of a producer, loaded one chunk ahead.
struct SySurfel {
    float cx, cy, cz, nx, ny, nz;
    double2 amp;
};

__device__ __forceinline__ SySurfel sy_load(int s, const float* __restrict__ ctr, const float* __restrict__ nrm,
                                            const double2* __restrict__ amp) {
    return SySurfel{ctr[3 * s], ctr[3 * s + 1], ctr[3 * s + 2], nrm[3 * s], nrm[3 * s + 1], nrm[3 * s + 2], amp[s]};
}

// Producer: the record of pair (aperture at p, surfel in) in float64, as groundtruth.py:69-84.
template <int J>
__device__ __forceinline__ void sy_produce(const Chirp& c, bool ok, double px, double py, double pz,
                                           const SySurfel& in, double* mine) {
    const double dx = px - (double)in.cx, dy = py - (double)in.cy, dz = pz - (double)in.cz;
    const double r = fmax(sqrt(fma(dx, dx, fma(dy, dy, dz * dz))), 1e-9);
    const double inv = 1.0 / r;
    const double cosine = fmax(dx * inv * (double)in.nx + dy * inv * (double)in.ny + dz * inv * (double)in.nz, 0.0);
    double gain = cosine * cosine * (c.lam_d / (8.0 * kPiD)) * inv;
    if (c.fresnel_d > 1.0) {
        const double root = sqrt(c.fresnel_d * c.fresnel_d - 1.0 + cosine * cosine);
        gain *= (root - cosine) / (root + cosine);
    }
    const double delay = c.two_over_c0 * r;
    const double2 rot = cis_cycles(c.fc * delay + c.slope * delay * (c.t0 - 0.5 * delay));
    double2 e = zmul(rot, in.amp);
    const double g = ok ? gain : 0.0;
    e = make_double2(g * e.x, g * e.y);
    const double f = r * c.k0pm * c.inv_N;  // cycles per ADC sample
    const double2 v = cis_cycles(f);
    const double2 z1 = cis_cycles(16.0 * (f - rint(f)));
    // Z_j = z1^j by a binary product tree
    double2 Z[J];
    Z[0] = make_double2(1.0, 0.0);
    Z[1] = z1;
#pragma unroll
    for (int j = 2; j < J; ++j) Z[j] = (j & 1) ? zmul(Z[j - 1], z1) : zsq(Z[j >> 1]);
#pragma unroll
    for (int j = 0; j < J; ++j) {
        mine[j] = Z[j].x;
        mine[J + j] = Z[j].y;
    }
    // e W_i = e v^i, two interleaved chains by v^2
    const double2 v2 = zsq(v);
    double2 w0 = e, w1 = zmul(e, v);
#pragma unroll
    for (int i = 0; i < 16; i += 2) {
        mine[2 * J + i] = w0.x;
        mine[2 * J + 16 + i] = w0.y;
        mine[2 * J + i + 1] = w1.x;
        mine[2 * J + 17 + i] = w1.y;
        w0 = zmul(w0, v2);
        w1 = zmul(w1, v2);
    }
}

// Consumer: wave cw accumulates its 4 apertures over one chunk on v_mfma_f64_16x16x4f64.
// Lane layouts (gfx950, measured): A[i][k] -> lane i + 16 k, B[k][j] -> lane j + 16 k, and
// C[lane / 16 + 4 r][lane % 16] in register r (rows interleaved, unlike the FP32 form).
// K index kk = surfel 4 t + kk; row x of tile tt = (Z part x / 8, j = 8 tt + x % 8), so
// registers 0, 1 of lane (g, i) hold Re-part rows j = g, g + 4 and registers 2, 3 the Im-part
// rows of the same j; column tile p = (e W) part. Re X = C0[(j, re)] - C1[(j, im)],
// Im X = C1[(j, re)] + C0[(j, im)].
template <int J>
__device__ __forceinline__ void sy_consume(const double* rec, int cw, int lane, f64x4 (&acc)[4][2][J / 8]) {
    constexpr int REC = SyRec<J>::size, T = J / 8;
    const int kk = lane >> 4, x = lane & 15;
    const int a_off = (x >> 3) * J + (x & 7);
    const int b_off = 2 * J + x;
#pragma unroll
    for (int t = 0; t < kSyChunk / 4; ++t)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const double* r = rec + ((4 * cw + b) * kSyChunk + 4 * t + kk) * REC;
            const double br = r[b_off], bi = r[b_off + 16];
#pragma unroll
            for (int tt = 0; tt < T; ++tt) {
                const double av = r[a_off + 8 * tt];
                acc[b][0][tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, br, acc[b][0][tt], 0, 0, 0);
                acc[b][1][tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bi, acc[b][1][tt], 0, 0, 0);
            }
        }
}

template <int J>
__global__ __launch_bounds__(128) void synthesize_kernel(const Chirp c, const float* __restrict__ pos, int A,
                                                         const float* __restrict__ ctr,
                                                         const float* __restrict__ nrm,
                                                         const double2* __restrict__ amp, int S,
                                                         double2* __restrict__ out) {
    constexpr int AT = kSyApertures, SC = kSyChunk, T = J / 8;
    extern __shared__ double dsm[];
    double2* E = reinterpret_cast<double2*>(dsm);  // [N]: e^{-2 pi i m / N}
    double* rec = dsm + 2 * c.N;                   // [AT * SC][REC]; later X [AT][16 J + 1]

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int a0 = blockIdx.x * AT;
    for (int m = tid; m < c.N; m += 128) {
        double s_, c_;
        sincospi(-2.0 * (double)m / (double)c.N, &s_, &c_);
        E[m] = make_double2(c_, s_);
    }
    const int pal = tid >> 4, psp = tid & 15;  // producer pair: aperture a0 + pal, surfel psp
    const int pa = min(a0 + pal, A - 1);
    const bool p_ok = a0 + pal < A;
    const double px = pos[3 * pa], py = pos[3 * pa + 1], pz = pos[3 * pa + 2];

    f64x4 acc[4][2][T];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int tt = 0; tt < T; ++tt) acc[b][p][tt] = f64x4{0.0, 0.0, 0.0, 0.0};
    SySurfel in = sy_load(min(psp, S - 1), ctr, nrm, amp);
    for (int s0 = 0; s0 < S; s0 += SC) {
        const int s = s0 + psp;
        const SySurfel next = sy_load(min(s + SC, S - 1), ctr, nrm, amp);
        __syncthreads();  // previous chunk consumed (and the twiddle table written)
        sy_produce<J>(c, p_ok && s < S, px, py, pz, in, rec + tid * SyRec<J>::size);
        __syncthreads();
        sy_consume<J>(rec, wave, lane, acc);
        in = next;
    }
    __syncthreads();
    // X rows to LDS: registers h and h + 2 of lane (g, i) in tile tt are the Re- and Im-part
    // rows of j = 8 tt + g + 4 h, sample 16 j + i
    constexpr int LDX = 16 * J + 1;
    double2* X = reinterpret_cast<double2*>(rec);
    {
        const int g = lane >> 4, i = lane & 15;
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int tt = 0; tt < T; ++tt)
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int j = 8 * tt + g + 4 * h;
                    X[(4 * wave + b) * LDX + 16 * j + i] = make_double2(acc[b][0][tt][h] - acc[b][1][tt][h + 2],
                                                                        acc[b][1][tt][h] + acc[b][0][tt][h + 2]);
                }
    }
    __syncthreads();
    // out[a0 + b][k] = (1/samples) sum_{n < samples} X[b][n] E[n k mod N]; consecutive threads
    // take consecutive bins of one aperture (coalesced stores)
    const int rows = min(AT, A - a0);
    const double scale = 1.0 / (double)c.S;
    for (int idx = tid; idx < rows * c.K; idx += 128) {
        const int b = idx / c.K, k = idx - b * c.K;
        const double2* xr = X + b * LDX;
        double re = 0.0, im = 0.0;
        int m = 0;
        for (int n = 0; n < c.S; ++n) {
            const double2 xv = xr[n], ev = E[m];
            re = fma(xv.x, ev.x, fma(-xv.y, ev.y, re));
            im = fma(xv.x, ev.y, fma(xv.y, ev.x, im));
            m += k;
            m -= m >= c.N ? c.N : 0;
        }
        out[(size_t)(a0 + b) * c.K + k] = make_double2(scale * re, scale * im);
    }
}

// The MFMA tiles cover n < 16 J (J = 8 or 16); the DFT reads the first `samples` of them.
bool synthesize_supported(const Chirp& c) { return c.S <= 256; }

template <int J>
static hipError_t launch_synthesize_j(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                                      const double2* amp, int S, double2* out, hipStream_t st) {
    const size_t rec = sizeof(double) * (size_t)kSyApertures * kSyChunk * SyRec<J>::size;
    const size_t xbuf = sizeof(double2) * (size_t)kSyApertures * (16 * J + 1);
    const size_t smem = sizeof(double2) * c.N + (rec > xbuf ? rec : xbuf);
    hipError_t e = allow_lds((const void*)synthesize_kernel<J>, smem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(synthesize_kernel<J>, dim3((A + kSyApertures - 1) / kSyApertures), dim3(128), smem, st, c, pos,
                       A, ctr, nrm, amp, S, out);
    return hipGetLastError();
}

hipError_t launch_synthesize(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                             const double2* amp, int S, double2* out, hipStream_t st) {
    if (!synthesize_supported(c)) return hipErrorInvalidValue;
    if (c.S <= 128) return launch_synthesize_j<8>(c, pos, A, ctr, nrm, amp, S, out, st);
    return launch_synthesize_j<16>(c, pos, A, ctr, nrm, amp, S, out, st);
}

}  // namespace dsplat
