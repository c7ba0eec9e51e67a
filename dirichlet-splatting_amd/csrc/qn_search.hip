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

// The damped quasi-Newton direction from the inverse-Hessian estimate H [6][6] and the ascent
// gradient g at normals n: -H g, or -g when that does not ascend; masked, the normal part
// tangent to n, shortened to the trust radius (search.slang:66-88).
template <int W>
__device__ __forceinline__ void qn_direction(const float* H, const float* g, const float* n, const QnSpec& sp,
                                             float* d) {
    float descent = 0.0f;
#pragma unroll
    for (int j = 0; j < W; ++j) {
        float v = 0.0f;
#pragma unroll
        for (int k = 0; k < W; ++k) v -= H[j * 6 + k] * g[k];
        d[j] = v;
        descent += v * g[j];
    }
#pragma unroll
    for (int j = 0; j < W; ++j) d[j] = (descent >= 0.0f ? -g[j] : d[j]) * sp.direction_mask[j];
    if constexpr (W == 6) {
        float proj = 0.0f;
#pragma unroll
        for (int j = 0; j < 3; ++j) proj += d[j + 3] * n[j];
#pragma unroll
        for (int j = 0; j < 3; ++j) d[j + 3] -= proj * n[j];
    }
    float len2 = 0.0f;
#pragma unroll
    for (int j = 0; j < W; ++j) len2 += d[j] * d[j];
    const float shrink = fminf(1.0f, sp.radius / fmaxf(sqrtf(len2), 1e-30f));
#pragma unroll
    for (int j = 0; j < W; ++j) d[j] *= shrink;
}

// The trial point at `fraction` of direction d from (oc, on): centres clamped to the box,
// normals renormalized (search.slang:90-109).
template <int W>
__device__ __forceinline__ void qn_trial(const float* oc, const float* on, const float* d, float fraction,
                                         const QnSpec& sp, float* tc, float* tn) {
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
}

// Step and gradient change of an accepted move in scaled coordinates, and whether the
// curvature condition s . y > 1e-8 |s| |y| (and > 1e-20) admits the BFGS update; rho = 1 / s . y.
template <int W>
__device__ __forceinline__ bool qn_secant(const QnSpec& sp, const float* c, const float* n, const float* oc,
                                          const float* on, const float* g_new, const float* g, float* step,
                                          float* change, float* rho) {
    float sy = 0.0f, ss = 0.0f, yy = 0.0f;
#pragma unroll
    for (int j = 0; j < W; ++j) {
        step[j] = j < 3 ? (c[j] - oc[j]) / sp.center_scale : (n[j - 3] - on[j - 3]) / sp.normal_scale;
        change[j] = g_new[j] - g[j];
        sy += step[j] * change[j];
        ss += step[j] * step[j];
        yy += change[j] * change[j];
    }
    *rho = 1.0f / fmaxf(sy, 1e-30f);
    return sy > 1e-8f * sqrtf(ss) * sqrtf(yy) && sy > 1e-20f;
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
        float d[6];
        qn_direction<W>(H, st.g, st.n, sp, d);

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
            qn_trial<W>(oc, on, d, fraction, sp, tc, tn);
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
        float step[6], change[6], rho;
        const bool valid = qn_secant<W>(sp, st.c, st.n, oc, on, g_new, st.g, step, change, &rho);
        if (valid) {  // uniform; H <- L H L^T + rho s s^T, L = I - rho s y^T, one element per thread
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


// ---------------------------------------------------------------------------------------
// Lockstep ascent on sample-domain chirps
// ---------------------------------------------------------------------------------------
// The same per-candidate ascent as qn_search_kernel, taken apart into rounds: one
// certificate_gradient launch evaluates the certificate sums and their derivatives (21
// values) of every live candidate at its query point on the matrix cores, then one thread
// per candidate moves its state machine:
//   round 0          the starting point's value and gradient; the first direction;
//   a trial rises    accept it: its gradient came with it, the BFGS update, the next direction;
//   a trial fails    halve the fraction, or stop after `backtracks` trials;
// and writes its next query point. An accepted trial costs one round (its value and gradient
// come from one evaluation), so an iteration whose full step rises takes one round where the
// one-launch kernel makes two passes over the apertures. Decisions compare values of the one
// kernel, as there.
struct QnLockState {
    float c[3], n[3];    // current point
    float oc[3], on[3];  // start of the line search
    float d[6];          // its direction
    float g[6];          // ascent gradient at the current point
    float H[36];         // inverse-Hessian estimate [6][6]
    float value, fraction;
    int backtrack, iteration;
};

__global__ __launch_bounds__(256) void qn_lock_init_kernel(const float* __restrict__ ctr, const float* __restrict__ nrm,
                                                           int C, float* __restrict__ qc, float* __restrict__ qn,
                                                           int* __restrict__ live, int* __restrict__ counters) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s == 0) counters[0] = counters[1] = 0;
    if (s >= C) return;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        qc[3 * s + j] = ctr[3 * s + j];
        qn[3 * s + j] = nrm[3 * s + j];
    }
    live[s] = 1;
}

// One round's step (first: round 0); counters[parity] counts the candidates still live after
// it, and the other counter is cleared for the next round.
template <int W>
__global__ __launch_bounds__(256) void qn_lock_step_kernel(const float* __restrict__ sums,
                                                           const float* __restrict__ areas, int C, const QnSpec sp,
                                                           QnLockState* __restrict__ states, float* __restrict__ qc,
                                                           float* __restrict__ qn, int* __restrict__ live,
                                                           int* __restrict__ counters, int first, int parity) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s == 0) counters[parity ^ 1] = 0;
    if (s >= C || live[s] == 0) return;
    QnLockState st = states[s];
    // certificate_gradient's columns (corr, norm, dcorr re [6], im [6], dnorm [6]) in the
    // layout of qn_certificate for W coordinates
    const float* v = sums + (size_t)s * kCgValues;
    float t[kQnMaxValues];
#pragma unroll
    for (int j = 0; j < 3; ++j) t[j] = v[j];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        t[3 + j] = v[3 + j];
        t[3 + W + j] = v[9 + j];
        t[3 + 2 * W + j] = v[15 + j];
    }
    float pc[3], pn[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        pc[j] = qc[3 * s + j];
        pn[j] = qn[3 * s + j];
    }
    float g[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const float value = qn_certificate<W>(sp, t, areas[s], pn, g);
    bool going, begin;
    if (first) {  // the starting point
        st.value = value;
#pragma unroll
        for (int j = 0; j < 6; ++j) st.g[j] = g[j];
#pragma unroll
        for (int i = 0; i < 36; ++i) st.H[i] = i / 6 == i % 6 ? 1.0f : 0.0f;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            st.c[j] = pc[j];
            st.n[j] = pn[j];
        }
        st.iteration = 0;
        going = begin = sp.iterations > 0 && sp.backtracks > 0;
    } else if (isfinite(value) && value > st.value) {  // the trial rises
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            st.c[j] = pc[j];
            st.n[j] = pn[j];
        }
        float step[6], change[6], rho;
        if (qn_secant<W>(sp, st.c, st.n, st.oc, st.on, g, st.g, step, change, &rho)) {
            // H <- L H L^T + rho s s^T, L = I - rho s y^T, in the order of qn_search_kernel
            float LH[36];
#pragma unroll
            for (int i = 0; i < W; ++i)
#pragma unroll
                for (int j = 0; j < W; ++j) {
                    float x = 0.0f;
#pragma unroll
                    for (int k = 0; k < W; ++k) x += ((i == k ? 1.0f : 0.0f) - rho * step[i] * change[k]) * st.H[k * 6 + j];
                    LH[i * 6 + j] = x;
                }
#pragma unroll
            for (int i = 0; i < W; ++i)
#pragma unroll
                for (int j = 0; j < W; ++j) {
                    float x = 0.0f;
#pragma unroll
                    for (int k = 0; k < W; ++k) x += LH[i * 6 + k] * ((j == k ? 1.0f : 0.0f) - rho * step[j] * change[k]);
                    st.H[i * 6 + j] = x + rho * step[i] * step[j];
                }
        }
        st.value = value;
#pragma unroll
        for (int j = 0; j < 6; ++j) st.g[j] = g[j];
        going = begin = ++st.iteration < sp.iterations;
    } else {  // the trial fails: a shorter one, or stop
        going = ++st.backtrack < sp.backtracks;
        begin = false;
        st.fraction *= 0.5f;
    }
    if (begin) {
        qn_direction<W>(st.H, st.g, st.n, sp, st.d);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            st.oc[j] = st.c[j];
            st.on[j] = st.n[j];
        }
        st.fraction = 1.0f;
        st.backtrack = 0;
    }
    if (going) {
        qn_trial<W>(st.oc, st.on, st.d, st.fraction, sp, pc, pn);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            qc[3 * s + j] = pc[j];
            qn[3 * s + j] = pn[j];
        }
        atomicAdd(&counters[parity], 1);
    } else {
        live[s] = 0;
    }
    states[s] = st;
}

__global__ __launch_bounds__(256) void qn_lock_finish_kernel(const QnLockState* __restrict__ states, int C,
                                                             float* __restrict__ ctr_out, float* __restrict__ nrm_out,
                                                             float* __restrict__ value_out) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= C) return;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        ctr_out[3 * s + j] = states[s].c[j];
        nrm_out[3 * s + j] = states[s].n[j];
    }
    value_out[s] = states[s].value;
}

struct QnLockLayout {
    size_t states, qc, qn, sums, live, counters, grad, total;  // byte offsets
};

static QnLockLayout qn_lock_layout(const Chirp& c, int A, int C) {
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    QnLockLayout o;
    o.states = 0;
    o.qc = up(sizeof(QnLockState) * (size_t)C);
    o.qn = o.qc + up(sizeof(float) * 3 * (size_t)C);
    o.sums = o.qn + up(sizeof(float) * 3 * (size_t)C);
    o.live = o.sums + up(sizeof(float) * kCgValues * (size_t)C);
    o.counters = o.live + up(sizeof(int) * (size_t)C);
    o.grad = o.counters + 256;
    o.total = o.grad + up(cert_grad_workspace(c, A, C));
    return o;
}

size_t qn_lockstep_workspace(const Chirp& c, int A, int count) {
    if (cert_grad_workspace(c, A, count) == 0) return 0;
    return qn_lock_layout(c, A, count).total;
}

constexpr int kQnLockCheck = 4;  // rounds between reads of the live count

hipError_t launch_qn_lockstep(const Chirp& c, const float* pos, int A, const void* operands, const float* ctr,
                              const float* nrm, const float* areas, int C, const QnSpec& sp, float* oc, float* on,
                              float* ov, void* workspace, hipStream_t st) {
    const QnLockLayout o = qn_lock_layout(c, A, C);
    char* base = static_cast<char*>(workspace);
    QnLockState* states = reinterpret_cast<QnLockState*>(base + o.states);
    float* qc = reinterpret_cast<float*>(base + o.qc);
    float* qn = reinterpret_cast<float*>(base + o.qn);
    float* sums = reinterpret_cast<float*>(base + o.sums);
    int* live = reinterpret_cast<int*>(base + o.live);
    int* counters = reinterpret_cast<int*>(base + o.counters);
    const dim3 grid((C + 255) / 256);
    hipLaunchKernelGGL(qn_lock_init_kernel, grid, dim3(256), 0, st, ctr, nrm, C, qc, qn, live, counters);
    hipError_t e = hipGetLastError();
    // round 0 evaluates the starting points; each later round one trial per live candidate,
    // at most `backtracks` per iteration
    const long long rounds = 1 + (long long)sp.iterations * sp.backtracks;
    for (long long r = 0; r < rounds && e == hipSuccess; ++r) {
        e = launch_cert_grad(c, pos, A, qc, qn, C, operands, sums, base + o.grad, st, live);
        if (e != hipSuccess) break;
        if (sp.width == 6)
            hipLaunchKernelGGL(qn_lock_step_kernel<6>, grid, dim3(256), 0, st, sums, areas, C, sp, states, qc, qn, live,
                               counters, (int)(r == 0), (int)(r & 1));
        else
            hipLaunchKernelGGL(qn_lock_step_kernel<3>, grid, dim3(256), 0, st, sums, areas, C, sp, states, qc, qn, live,
                               counters, (int)(r == 0), (int)(r & 1));
        if ((e = hipGetLastError()) != hipSuccess) break;
        if (r % kQnLockCheck == kQnLockCheck - 1 && r + 1 < rounds) {
            int left = 0;
            e = hipMemcpyAsync(&left, counters + (r & 1), sizeof(int), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess || left == 0) break;
        }
    }
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(qn_lock_finish_kernel, grid, dim3(256), 0, st, states, C, oc, on, ov);
    return hipGetLastError();
}

}  // namespace dsplat
