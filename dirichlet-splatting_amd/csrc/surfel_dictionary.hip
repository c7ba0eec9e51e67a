// Dictionary atoms and the explicit Jacobian (reference: dsplat/kernels/surfel.slang:313-378
// design_bins; correlate.slang:237-281 surfel_design_jacobian; bindings
// dsplat/kernels/surfel.py:199-239).
#include "dsplat_device.h"
#include "dsplat_internal.h"

namespace dsplat {

// ---------------------------------------------------------------------------------------
// Design: entries c sqrt(area) D(k - k0). A block is 64 surfels x 4 apertures (one per
// wave); each lane owns one pair and walks the bins, so for every bin a wave stores 64
// consecutive surfels: 512 B (complex) or 256 B (real) per store, coalesced. Layouts:
//   complex  out[(a K + k) S + s] = entry                       ([A K, S] c64)
//   real     out[((2 k + part) A + a) S + s] = re / im           ([2 K A, S] f32)
// HBM-write bound: 8 B per entry.
// ---------------------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(256) void design_kernel(const Chirp c, const float* __restrict__ pos, int A,
                                                     const float* __restrict__ ctr, const float* __restrict__ nrm,
                                                     const float* __restrict__ areas, int S, int complex_layout,
                                                     float* __restrict__ out) {
    extern __shared__ float T[];  // [2N], T_0 = 0
    const int tid = threadIdx.x;
    for (int i = tid; i < 2 * c.N; i += 256) T[i] = cot_entry(i, c.N, 0.0f);
    __syncthreads();
    const int s = blockIdx.x * 64 + (tid & 63);
    const int a = blockIdx.y * 4 + (tid >> 6);
    if (a >= A) return;  // whole wave
    const bool ok = s < S;
    const int sv = ok ? s : S - 1;
    const Pair q = pair_geometry(c, pos[3 * a], pos[3 * a + 1], pos[3 * a + 2], ctr[3 * sv], ctr[3 * sv + 1],
                                 ctr[3 * sv + 2], nrm[3 * sv], nrm[3 * sv + 1], nrm[3 * sv + 2]);
    const float scale = q.amp * sqrtf(areas[sv]);
    const float2 cw = scale * q.rot;
    const float t = tan_small(c.pi_N * q.delta);
    float sb[P], cb[P];
    class_angles<P>(q.delta, q.lobe, sb, cb);
    float2 g[P];
#pragma unroll
    for (int r = 0; r < P; ++r) g[r] = c.inv_S * cmul(cw, make_float2(sb[r] * cb[r], -sb[r] * sb[r]));
    float2 D, Dp;
    lobe_direct(c, q.delta, &D, &Dp);
    const float2 lobe_value = cmul(cw, D);
    const float* Tb = T + (c.N - q.lobe);
    for (int k0 = 0; k0 < c.K; k0 += 16) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int k = k0 + j;
            if (k < c.K) {  // uniform across the wave
                const float cot = cot_bin(Tb[k], t);
                const float2 gr = g[j % P];
                float2 v = make_float2(gr.x * cot - gr.y, gr.y * cot + gr.x);
                if (k == q.lobe) v = lobe_value;
                if (ok) {
                    if (complex_layout) {
                        reinterpret_cast<float2*>(out)[((size_t)a * c.K + k) * S + s] = v;
                    } else {
                        out[((size_t)(2 * k) * A + a) * S + s] = v.x;
                        out[((size_t)(2 * k + 1) * A + a) * S + s] = v.y;
                    }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// Jacobian: one thread per entry (a, k, s), s fastest. Columns [A,K,S] c64 and their
// derivatives by center xyz and normal xyz [A,K,S,6] c64, without area:
//   d col / d c_j = rot [da_j D - i amp rate u_j D + amp k0pm u_j D'],
//   d col / d n_j = rot amp_cos u_j D.
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void jacobian_kernel(const Chirp c, const float* __restrict__ pos, int A,
                                                       const float* __restrict__ ctr,
                                                       const float* __restrict__ nrm, int S,
                                                       float2* __restrict__ col, float2* __restrict__ jac) {
    const size_t i = blockIdx.x * (size_t)256 + threadIdx.x;
    const size_t total = (size_t)A * c.K * S;
    if (i >= total) return;
    const int s = (int)(i % S);
    const int k = (int)((i / S) % c.K);
    const int a = (int)(i / ((size_t)S * c.K));
    const float nx = nrm[3 * s], ny = nrm[3 * s + 1], nz = nrm[3 * s + 2];
    const Pair q = pair_geometry(c, pos[3 * a], pos[3 * a + 1], pos[3 * a + 2], ctr[3 * s], ctr[3 * s + 1],
                                 ctr[3 * s + 2], nx, ny, nz);
    float2 value = make_float2(0.f, 0.f);
    float2 d[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) d[j] = make_float2(0.f, 0.f);
    if (q.amp != 0.0f) {
        float2 D, Dp;
        kernel_at(c, k - q.lobe, q.delta, &D, &Dp);
        const float2 z = cmul(D, q.rot), dz = cmul(Dp, q.rot);
        value = q.amp * z;
        float da[3];
        amp_gradient(q, nx, ny, nz, da);
        const float u[3] = {q.ux, q.uy, q.uz};
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float2 iz = make_float2(-z.y, z.x);
            d[j] = da[j] * z + q.amp * u[j] * (c.k0pm_f * dz - q.rate * iz);
            d[j + 3] = (q.amp_cos * u[j]) * z;
        }
    }
    col[i] = value;
#pragma unroll
    for (int j = 0; j < 6; ++j) jac[i * 6 + j] = d[j];
}

hipError_t launch_design(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                         const float* areas, int S, int complex_layout, float* out, hipStream_t st) {
    const dim3 grid((S + 63) / 64, (A + 3) / 4);
    const size_t smem = sizeof(float) * 2 * c.N;
    switch (c.P) {
        case 1: hipLaunchKernelGGL(design_kernel<1>, grid, dim3(256), smem, st, c, pos, A, ctr, nrm, areas, S, complex_layout, out); break;
        case 2: hipLaunchKernelGGL(design_kernel<2>, grid, dim3(256), smem, st, c, pos, A, ctr, nrm, areas, S, complex_layout, out); break;
        case 4: hipLaunchKernelGGL(design_kernel<4>, grid, dim3(256), smem, st, c, pos, A, ctr, nrm, areas, S, complex_layout, out); break;
        case 8: hipLaunchKernelGGL(design_kernel<8>, grid, dim3(256), smem, st, c, pos, A, ctr, nrm, areas, S, complex_layout, out); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_jacobian(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm, int S,
                           float2* col, float2* jac, hipStream_t st) {
    const size_t total = (size_t)A * c.K * S;
    if (total == 0) return hipSuccess;
    const size_t blocks = (total + 255) / 256;
    if (blocks > 0x7fffffffu) return hipErrorInvalidValue;
    hipLaunchKernelGGL(jacobian_kernel, dim3((unsigned)blocks), dim3(256), 0, st, c, pos, A, ctr, nrm, S, col, jac);
    return hipGetLastError();
}

}  // namespace dsplat
