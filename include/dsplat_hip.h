/*
 * C ABI of the MI355X (gfx950) Dirichlet-splat library, libdsplat_hip.so.
 *
 * Every pointer is a device pointer to contiguous row-major float32 data unless
 * stated otherwise; complex values are interleaved (re, im) float pairs. Every
 * entry point enqueues its kernels on `stream` (a hipStream_t; NULL is the
 * default stream), allocates nothing, never throws, and returns 0 (hipSuccess)
 * or a hipError_t code: hipErrorInvalidValue (1) for rejected shapes or
 * configurations, otherwise the launch error. dsplat_error_string() names a code.
 *
 * The reference binds each of these operations from Python (slangtorch launches
 * in the dsplat/kernels Python modules); the cited file:line is the binding each entry point
 * replaces. Outputs and workspaces are allocated by the caller, as the reference
 * bindings allocate them through torch.
 */
#ifndef DSPLAT_HIP_H
#define DSPLAT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* FMCW chirp constants (dsplat/fmcw.py:29-65, sensor_args at kernels/surfel.py:21-32).
 * num_bins = n_fft / 2, n_fft = samples * pad, pad in {1, 2, 4, 8}, num_bins <= 512. */
typedef struct dsplat_fmcw {
    int32_t num_bins;   /* K, observed range bins */
    int32_t n_fft;      /* N, FFT length */
    int32_t samples;    /* S, ADC samples */
    int32_t pad;        /* P = N / S */
    double k0_per_m;    /* 1 / range_bin_spacing_m */
    double fc_hz;       /* carrier */
    double slope_hz_s;  /* chirp slope */
    double t_start_s;   /* ADC start time */
    double lambda_m;    /* c0 / fc */
    double c0_m_s;      /* wave speed */
    double fresnel_n;   /* refractive index; <= 1 disables the Fresnel factor */
} dsplat_fmcw;

/* Library version and a readable name for a returned code. */
const char* dsplat_version(void);
const char* dsplat_error_string(int code);

/* Range profiles out[A, K] = sum_s w_s column_s(p_a).
 * positions [A,3], centers [S,3], normals [S,3], weights [S] complex, out [A,K] complex.
 * Replaces kernels/surfel.py:58-108 (_profiles: pair_geometry + render_bins). */
int dsplat_surfel_render(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                         const float* centers, const float* normals, const float* weights,
                         int32_t num_surfels, float* out, void* stream);

/* Workspace bytes dsplat_surfel_vjp needs for A apertures and S surfels. */
size_t dsplat_surfel_vjp_workspace(int32_t num_apertures, int32_t num_surfels);

/* Gradients of Re(conj(upstream) . render) for centers [S,3], normals [S,3] and the
 * complex weights [S] (PyTorch convention: grad_weights = A^H upstream).
 * upstream [A,K] complex. Deterministic: per-block partials reduced in float64.
 * Replaces kernels/surfel.py:123-137 (vjp: vjp_bins + host double sum). */
int dsplat_surfel_vjp(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                      const float* centers, const float* normals, const float* weights,
                      int32_t num_surfels, const float* upstream, float* grad_centers,
                      float* grad_normals, float* grad_weights, void* workspace,
                      size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DSPLAT_HIP_H */
