// Candidate kernels of the spatial search: range gate, greedy separated peaks and the
// batched quasi-Newton direction and trial step (reference: dsplat/kernels/search.slang:1-109,
// bindings dsplat/kernels/search.py:1-77).
#include <float.h>

#include "dsplat_internal.h"

namespace dsplat {

// A point passes when its spherical range bin (round half to even) falls in a marked
// interval at `required` apertures or more (search.slang:7-24).
__global__ __launch_bounds__(256) void range_gate_kernel(const float* __restrict__ pts, const float* __restrict__ pos,
                                                         const int* __restrict__ intervals, int count, int apertures,
                                                         int bins, float spacing, int required,
                                                         int* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
    int matches = 0;
    for (int a = 0; a < apertures; ++a) {
        const float dx = px - pos[3 * a], dy = py - pos[3 * a + 1], dz = pz - pos[3 * a + 2];
        const float x = __fsqrt_rn(__fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz))) / spacing;
        const float lo = floorf(x);
        int k = (int)lo;
        const float frac = x - lo;
        if (frac > 0.5f || (frac == 0.5f && (k & 1))) ++k;
        if (k >= 0 && k < bins && intervals[(size_t)a * bins + k] != 0) ++matches;
    }
    out[i] = matches >= required ? 1 : 0;
}

// Greedy best-first separated peaks in one workgroup; scores is a scratch copy the
// kernel overwrites (search.slang:27-64). Ties go to the lower index.
__global__ __launch_bounds__(256) void peaks_kernel(float* __restrict__ scores, const float* __restrict__ ctr,
                                                    int count, int peaks, int dim, float separation,
                                                    int* __restrict__ out) {
    __shared__ float best_v[256];
    __shared__ int best_i[256];
    const int lane = threadIdx.x;
    const float sep2 = separation * separation;
    for (int p = 0; p < peaks; ++p) {
        float best = -INFINITY;
        int index = -1;
        for (int i = lane; i < count; i += 256) {
            const float v = scores[i];
            if (isfinite(v) && (v > best || index < 0)) {  // ascending i: first max wins
                best = v;
                index = i;
            }
        }
        best_v[lane] = best;
        best_i[lane] = index;
        __syncthreads();
        for (int stride = 128; stride > 0; stride >>= 1) {
            if (lane < stride) {
                const float v = best_v[lane + stride];
                const int o = best_i[lane + stride];
                if (o >= 0 && (best_i[lane] < 0 || v > best_v[lane] || (v == best_v[lane] && o < best_i[lane]))) {
                    best_v[lane] = v;
                    best_i[lane] = o;
                }
            }
            __syncthreads();
        }
        index = best_i[0];
        if (lane == 0) out[p] = index;
        if (index < 0) return;  // uniform: every lane read the same best_i[0]
        for (int i = lane; i < count; i += 256) {
            float d2 = 0.0f;
            for (int j = 0; j < dim; ++j) {
                const float d = ctr[(size_t)i * dim + j] - ctr[(size_t)index * dim + j];
                d2 = fmaf(d, d, d2);
            }
            if (i == index || d2 < sep2) scores[i] = -INFINITY;
        }
        __syncthreads();
    }
}

// -H^{-1} g (or -g if not a descent direction), masked, tangent to unit normals and
// scaled into the trust radius (search.slang:66-91).
__global__ __launch_bounds__(128) void direction_kernel(const float* __restrict__ inv, const float* __restrict__ grad,
                                                        const float* __restrict__ nrm, const float* __restrict__ mask,
                                                        int count, int width, float radius, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float d[6];
    float descent = 0.0f;
    for (int j = 0; j < width; ++j) {
        float v = 0.0f;
        for (int k = 0; k < width; ++k) v -= inv[((size_t)i * width + j) * width + k] * grad[(size_t)i * width + k];
        d[j] = v;
        descent += v * grad[(size_t)i * width + j];
    }
    for (int j = 0; j < width; ++j) d[j] = (descent >= 0.0f ? -grad[(size_t)i * width + j] : d[j]) * mask[j];
    if (width == 6) {
        float proj = 0.0f;
        for (int j = 0; j < 3; ++j) proj += d[j + 3] * nrm[3 * i + j];
        for (int j = 0; j < 3; ++j) d[j + 3] -= proj * nrm[3 * i + j];
    }
    float len2 = 0.0f;
    for (int j = 0; j < width; ++j) len2 += d[j] * d[j];
    const float scale = fminf(1.0f, radius / fmaxf(sqrtf(len2), 1e-30f));
    for (int j = 0; j < width; ++j) out[(size_t)i * width + j] = d[j] * scale;
}

// Step a fraction along the direction: centers clamped to the box where trainable,
// normals renormalized when they move (search.slang:93-109). controls [4, 6]:
// per-coordinate scale, trainable mask, lower and upper bounds.
__global__ __launch_bounds__(128) void trial_kernel(const float* __restrict__ ctr, const float* __restrict__ nrm,
                                                    const float* __restrict__ dir, const float* __restrict__ ctl,
                                                    int count, int width, float fraction, float* __restrict__ oc,
                                                    float* __restrict__ on) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float n[3];
    float len2 = 0.0f;
    for (int j = 0; j < 3; ++j) {
        const float c = ctr[3 * i + j] + fraction * ctl[j] * dir[(size_t)i * width + j];
        oc[3 * i + j] = ctl[6 + j] > 0.0f ? fminf(fmaxf(c, ctl[12 + j]), ctl[18 + j]) : ctr[3 * i + j];
        n[j] = nrm[3 * i + j] + (width == 6 ? fraction * ctl[j + 3] * dir[(size_t)i * width + j + 3] : 0.0f);
        len2 += n[j] * n[j];
    }
    const float inv = width == 6 ? 1.0f / sqrtf(len2) : 1.0f;
    for (int j = 0; j < 3; ++j) on[3 * i + j] = width == 6 ? n[j] * inv : nrm[3 * i + j];
}

hipError_t launch_range_gate(const float* pts, const float* pos, const int* intervals, int count, int apertures,
                             int bins, float spacing, int required, int* out, hipStream_t st) {
    hipLaunchKernelGGL(range_gate_kernel, dim3((count + 255) / 256), dim3(256), 0, st, pts, pos, intervals, count,
                       apertures, bins, spacing, required, out);
    return hipGetLastError();
}

hipError_t launch_peaks(float* scores, const float* ctr, int count, int peaks, int dim, float separation, int* out,
                        hipStream_t st) {
    hipLaunchKernelGGL(peaks_kernel, dim3(1), dim3(256), 0, st, scores, ctr, count, peaks, dim, separation, out);
    return hipGetLastError();
}

hipError_t launch_direction(const float* inv, const float* grad, const float* nrm, const float* mask, int count,
                            int width, float radius, float* out, hipStream_t st) {
    hipLaunchKernelGGL(direction_kernel, dim3((count + 127) / 128), dim3(128), 0, st, inv, grad, nrm, mask, count,
                       width, radius, out);
    return hipGetLastError();
}

hipError_t launch_trial(const float* ctr, const float* nrm, const float* dir, const float* ctl, int count, int width,
                        float fraction, float* oc, float* on, hipStream_t st) {
    hipLaunchKernelGGL(trial_kernel, dim3((count + 127) / 128), dim3(128), 0, st, ctr, nrm, dir, ctl, count, width,
                       fraction, oc, on);
    return hipGetLastError();
}

}  // namespace dsplat
