// Batched quasi-Newton ascent of the residual certificate, every candidate in one launch
// (reference: the host loop of certificate.py:106-201, search_certificate, around the
// direction and trial kernels of search.slang:66-109 and the correlations of
// correlate.slang:78-235).
//
// One workgroup per candidate. Its 256 threads split the apertures for every evaluation of
// the certificate (and, after an accepted step, its gradient), reduce the sums in float64,
// and then all hold the same candidate state: centers, normals, ascent gradient, the
// inverse-Hessian estimate and the current value. Per iteration: the damped quasi-Newton
// direction, then fractions 1, 1/2, 1/4, ... of it until the certificate rises, then the
// gradient at the new point and the BFGS update of the inverse. A candidate whose step
// finds no rise stops, which is what the host loop reaches as well: its state no longer
// changes while others continue.
//
// The certificate is |s|^2 / (max(norm, 1e-60) power), s = sqrt(area) a^H r (or its real
// part), norm = area |a|^2, power = |r|^2, formed in float64 from the float32-rounded
// sums like certificate.py:11-31; gradients by centers (and normals) likewise.
#include "dsplat_internal.h"
#include "dsplat_sums.h"

namespace dsplat {

constexpr int kQnThreads = 256;
constexpr int kQnMaxValues = 21;  // corr (2), norm, dcorr (2 x 6), dnorm (6)

// Aperture-minor copies for the per-thread aperture loops of qn_sums, so that the lanes of
// a wave (consecutive apertures) read consecutive addresses: resT[k][a] = res[a][k] and the
// class sums usT[r][a] = sum over bins k = r mod P of res[a][k].
__global__ __launch_bounds__(256) void qn_layout_kernel(const float2* __restrict__ res, int A, int K, int P,
                                                        float2* __restrict__ resT, float2* __restrict__ usT) {
    const int a = blockIdx.x * 256 + threadIdx.x;
    if (a >= A) return;
    const int k0 = blockIdx.y * 16;
    for (int k = k0; k < min(K, k0 + 16); ++k) resT[(size_t)k * A + a] = res[(size_t)a * K + k];
    if (blockIdx.y == 0) {
        for (int r = 0; r < P; ++r) {
            float2 s = make_float2(0.f, 0.f);
            for (int k = r; k < K; k += P) s += res[(size_t)a * K + k];
            usT[(size_t)r * A + a] = s;
        }
    }
}

// Per-thread sums over apertures tid, tid + 256, ... then a float64 workgroup reduction;
// every thread receives the totals. W = 0: corr, norm; W = 3 or 6 adds the derivatives.
template <int P, int W>
__device__ __forceinline__ void qn_sums(const Chirp& c, const float* __restrict__ T, const float* __restrict__ pos,
                                        int A, const float2* __restrict__ res, const float2* __restrict__ us,
                                        const float* cen, const float* nor, double* red, float* total) {
    constexpr int NV = 3 + 3 * W;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    float v[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] = 0.0f;
    for (int a = tid; a < A; a += kQnThreads) {
        const Pair q = pair_geometry(c, pos[3 * a], pos[3 * a + 1], pos[3 * a + 2], cen[0], cen[1], cen[2], nor[0],
                                     nor[1], nor[2]);
        const Sums z = bin_sums<P, (W > 0)>(c, T, res + a, us + a, q, A);
        const float2 corr = cmulc(q.rot, z.g0);
        v[0] += q.amp * corr.x;
        v[1] += q.amp * corr.y;
        v[2] += q.amp * q.amp * z.nrm;
        if constexpr (W > 0) {
            const float2 corr_x = cmulc(q.rot, z.g1);
            float da[3];
            amp_gradient(q, nor[0], nor[1], nor[2], da);
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
                v[3 + j] += dc.x;
                v[3 + W + j] += dc.y;
                v[3 + 2 * W + j] += dn;
            }
        }
    }
    __syncthreads();  // red is free: every thread has read the previous totals
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        double d = v[j];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) d += __shfl_xor(d, o, 64);
        if (lane == 0) red[wave * kQnMaxValues + j] = d;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        double d = 0.0;
        for (int w = 0; w < kQnThreads / 64; ++w) d += red[w * kQnMaxValues + j];
        total[j] = (float)d;
    }
}

struct QnState {
    float c[3], n[3];
    float g[6];  // ascent gradient (negated, masked, scaled, tangent)
    float value;
};

// The certificate from the rounded sums: value (float), and with W > 0 the ascent gradient.
template <int W>
__device__ __forceinline__ float qn_certificate(const QnSpec& sp, const float* t, float area, const float* n,
                                                float* g) {
    const float scale = sqrtf(area);
    const double cr = (double)(t[0] * scale), ci = (double)(t[1] * scale);
    const double nn = fmax((double)(t[2] * area), 1e-60);
    const double sr = cr, si = sp.real_part ? 0.0 : ci;
    const double mag = sp.real_part ? fabs(sr) : hypot(sr, si);
    const double value = mag * mag / nn / sp.power;
    if constexpr (W > 0) {
        float grad[6];
#pragma unroll
        for (int j = 0; j < W; ++j) {
            const double dcr = (double)(t[3 + j] * scale), dci = (double)(t[3 + W + j] * scale);
            const double dn = nn > 1e-60 ? (double)(t[3 + 2 * W + j] * area) : 0.0;
            const double ds = sp.real_part ? 2.0 * sr * dcr : 2.0 * (cr * dcr + ci * dci);
            grad[j] = (float)((ds - value * sp.power * dn) / nn / sp.power);
        }
        // ascent: -(tangent part * mask * scale) per field
#pragma unroll
        for (int j = 0; j < 3; ++j) g[j] = -(grad[j] * sp.center_mask[j] * sp.center_scale);
        if constexpr (W == 6) {
            float vec[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) vec[j] = grad[3 + j];
            if (sp.normalize) {
                float fixed2 = 0.0f, dot = 0.0f;
                float free_[3];
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const bool m = sp.normal_mask[j] > 0.0f;
                    free_[j] = m ? n[j] : 0.0f;
                    fixed2 += m ? 0.0f : n[j] * n[j];
                    vec[j] = m ? vec[j] : 0.0f;
                }
                const float room = fmaxf(1.0f - fixed2, 0.0f);
#pragma unroll
                for (int j = 0; j < 3; ++j) dot += vec[j] * free_[j];
#pragma unroll
                for (int j = 0; j < 3; ++j) vec[j] = room > 0.0f ? vec[j] - dot * free_[j] / fmaxf(room, 1e-30f) : 0.0f;
            }
#pragma unroll
            for (int j = 0; j < 3; ++j) g[3 + j] = -(vec[j] * sp.normal_mask[j] * sp.normal_scale);
        }
    }
    return (float)value;
}

template <int P, int W>
__global__ __launch_bounds__(kQnThreads) void qn_search_kernel(const Chirp c, const float* __restrict__ pos, int A,
                                                               const float2* __restrict__ res,
                                                               const float2* __restrict__ us,
                                                               const float* __restrict__ ctr0,
                                                               const float* __restrict__ nrm0,
                                                               const float* __restrict__ areas, const QnSpec sp,
                                                               float* __restrict__ ctr_out,
                                                               float* __restrict__ nrm_out,
                                                               float* __restrict__ value_out) {
    extern __shared__ float T[];  // [2N] derivative cot table
    __shared__ double red[(kQnThreads / 64) * kQnMaxValues];
    __shared__ float H[36], LH[36];  // inverse-Hessian estimate [6][6] and a product buffer
    const int s = blockIdx.x, tid = threadIdx.x;
    for (int i = tid; i < 2 * c.N; i += kQnThreads) T[i] = cot_entry(i, c.N, 0.0f);
    const float area = areas[s];
    QnState st;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        st.c[j] = ctr0[3 * s + j];
        st.n[j] = nrm0[3 * s + j];
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) st.g[i] = 0.0f;
    if (tid < 36) H[tid] = tid / 6 == tid % 6 ? 1.0f : 0.0f;
    __syncthreads();
    float t[kQnMaxValues];
    qn_sums<P, W>(c, T, pos, A, res, us, st.c, st.n, red, t);
    st.value = qn_certificate<W>(sp, t, area, st.n, st.g);

    for (int it = 0; it < sp.iterations; ++it) {
        // direction: -H g, or -g when that does not ascend; masked, tangent, within the radius
        float d[6];
        float descent = 0.0f;
#pragma unroll
        for (int j = 0; j < W; ++j) {
            float v = 0.0f;
#pragma unroll
            for (int k = 0; k < W; ++k) v -= H[j * 6 + k] * st.g[k];
            d[j] = v;
            descent += v * st.g[j];
        }
#pragma unroll
        for (int j = 0; j < W; ++j) d[j] = (descent >= 0.0f ? -st.g[j] : d[j]) * sp.direction_mask[j];
        if constexpr (W == 6) {
            float proj = 0.0f;
#pragma unroll
            for (int j = 0; j < 3; ++j) proj += d[j + 3] * st.n[j];
#pragma unroll
            for (int j = 0; j < 3; ++j) d[j + 3] -= proj * st.n[j];
        }
        float len2 = 0.0f;
#pragma unroll
        for (int j = 0; j < W; ++j) len2 += d[j] * d[j];
        const float shrink = fminf(1.0f, sp.radius / fmaxf(sqrtf(len2), 1e-30f));
#pragma unroll
        for (int j = 0; j < W; ++j) d[j] *= shrink;

        // backtracking along the direction
        float oc[3], on[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            oc[j] = st.c[j];
            on[j] = st.n[j];
        }
        bool accepted = false;
        float fraction = 1.0f;
        for (int b = 0; b < sp.backtracks; ++b, fraction *= 0.5f) {
            float tc[3], tn[3];
            float nl2 = 0.0f;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const float x = oc[j] + fraction * sp.step_scale[j] * d[j];
                tc[j] = sp.direction_mask[j] > 0.0f ? fminf(fmaxf(x, sp.lower[j]), sp.upper[j]) : oc[j];
                tn[j] = on[j] + (W == 6 ? fraction * sp.step_scale[j + 3] * d[j + 3] : 0.0f);
                nl2 += tn[j] * tn[j];
            }
            const float inv = W == 6 ? 1.0f / sqrtf(nl2) : 1.0f;
#pragma unroll
            for (int j = 0; j < 3; ++j) tn[j] = W == 6 ? tn[j] * inv : on[j];
            qn_sums<P, 0>(c, T, pos, A, res, us, tc, tn, red, t);
            float unused[6];
            const float trial = qn_certificate<0>(sp, t, area, tn, unused);
            if (isfinite(trial) && trial > st.value) {
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    st.c[j] = tc[j];
                    st.n[j] = tn[j];
                }
                accepted = true;
                break;
            }
        }
        if (!accepted) break;  // uniform: every thread holds the same state

        float g_new[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        qn_sums<P, W>(c, T, pos, A, res, us, st.c, st.n, red, t);
        const float v_new = qn_certificate<W>(sp, t, area, st.n, g_new);

        // BFGS update of the inverse with step and gradient change in scaled coordinates
        float step[6], change[6];
        float sy = 0.0f, ss = 0.0f, yy = 0.0f;
#pragma unroll
        for (int j = 0; j < W; ++j) {
            step[j] = j < 3 ? (st.c[j] - oc[j]) / sp.center_scale : (st.n[j - 3] - on[j - 3]) / sp.normal_scale;
            change[j] = g_new[j] - st.g[j];
            sy += step[j] * change[j];
            ss += step[j] * step[j];
            yy += change[j] * change[j];
        }
        const bool valid = sy > 1e-8f * sqrtf(ss) * sqrtf(yy) && sy > 1e-20f;
        if (valid) {  // uniform; H <- L H L^T + rho s s^T, L = I - rho s y^T, one element per thread
            const float rho = 1.0f / fmaxf(sy, 1e-30f);
            const int i = tid / W, j = tid - (tid / W) * W;
            if (tid < W * W) {
                float v = 0.0f;
                for (int k = 0; k < W; ++k) v += ((i == k ? 1.0f : 0.0f) - rho * step[i] * change[k]) * H[k * 6 + j];
                LH[i * 6 + j] = v;
            }
            __syncthreads();
            if (tid < W * W) {
                float v = 0.0f;
                for (int k = 0; k < W; ++k) v += LH[i * 6 + k] * ((j == k ? 1.0f : 0.0f) - rho * step[j] * change[k]);
                H[i * 6 + j] = v + rho * step[i] * step[j];
            }
            __syncthreads();
        }
        st.value = v_new;
#pragma unroll
        for (int j = 0; j < W; ++j) st.g[j] = g_new[j];
    }
    if (tid == 0) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            ctr_out[3 * s + j] = st.c[j];
            nrm_out[3 * s + j] = st.n[j];
        }
        value_out[s] = st.value;
    }
}

size_t qn_workspace(int A, int P, int K) { return sizeof(float2) * (size_t)A * (P + K); }

template <int P, int W>
static hipError_t launch_qn_pw(const Chirp& c, const float* pos, int A, const float2* res, const float2* us,
                               const float* ctr, const float* nrm, const float* areas, int count, const QnSpec& sp,
                               float* oc, float* on, float* ov, hipStream_t st) {
    const size_t smem = sizeof(float) * 2 * c.N;
    hipLaunchKernelGGL((qn_search_kernel<P, W>), dim3(count), dim3(kQnThreads), smem, st, c, pos, A, res, us, ctr,
                       nrm, areas, sp, oc, on, ov);
    return hipGetLastError();
}

template <int P>
static hipError_t launch_qn_p(const Chirp& c, const float* pos, int A, const float2* res, const float2* us,
                              const float* ctr, const float* nrm, const float* areas, int count, const QnSpec& sp,
                              float* oc, float* on, float* ov, hipStream_t st) {
    return sp.width == 6 ? launch_qn_pw<P, 6>(c, pos, A, res, us, ctr, nrm, areas, count, sp, oc, on, ov, st)
                         : launch_qn_pw<P, 3>(c, pos, A, res, us, ctr, nrm, areas, count, sp, oc, on, ov, st);
}

hipError_t launch_qn_search(const Chirp& c, const float* pos, int A, const float2* res, const float* ctr,
                            const float* nrm, const float* areas, int count, const QnSpec& sp, float* oc, float* on,
                            float* ov, void* workspace, hipStream_t st) {
    float2* us = static_cast<float2*>(workspace);  // [P][A]
    float2* resT = us + (size_t)c.P * A;            // [K][A]
    hipLaunchKernelGGL(qn_layout_kernel, dim3((A + 255) / 256, (c.K + 15) / 16), dim3(256), 0, st, res, A, c.K, c.P,
                       resT, us);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    res = resT;
    switch (c.P) {
        case 1: return launch_qn_p<1>(c, pos, A, res, us, ctr, nrm, areas, count, sp, oc, on, ov, st);
        case 2: return launch_qn_p<2>(c, pos, A, res, us, ctr, nrm, areas, count, sp, oc, on, ov, st);
        case 4: return launch_qn_p<4>(c, pos, A, res, us, ctr, nrm, areas, count, sp, oc, on, ov, st);
        case 8: return launch_qn_p<8>(c, pos, A, res, us, ctr, nrm, areas, count, sp, oc, on, ov, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace dsplat
