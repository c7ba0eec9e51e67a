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

/* Directional derivative out[A, K] of the profiles along tangents of the centers [S,3],
 * normals [S,3] and complex weights [S]. Replaces kernels/surfel.py:140-164 (jvp: jvp_bins). */
int dsplat_surfel_jvp(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                      const float* centers, const float* normals, const float* weights,
                      const float* delta_centers, const float* delta_normals, const float* delta_weights,
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

/* Dictionary rows of whole apertures including sqrt(area): complex [A*K, S]
 * (row a*K + k) when real_layout == 0, else real [2*K*A, S] with row (2k + part)*A + a.
 * areas [S]. Replaces kernels/surfel.py:199-220 (design: design_bins). */
int dsplat_surfel_design(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                         const float* centers, const float* normals, const float* areas,
                         int32_t num_surfels, int32_t real_layout, float* out, void* stream);

/* Explicit columns [A,K,S] complex and their derivatives by center xyz and normal xyz
 * [A,K,S,6] complex, without area. Replaces kernels/surfel.py:223-239
 * (jacobian: surfel_design_jacobian). */
int dsplat_surfel_jacobian(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                           const float* centers, const float* normals, int32_t num_surfels,
                           float* columns, float* derivatives, void* stream);

/* Workspace bytes of the correlation entry points: width = 3 (correlate),
 * 3 + 3w (gradient, w = 3 or 6) or 3 (1 + B) (orientations). */
size_t dsplat_surfel_correlate_workspace(int32_t num_apertures, int32_t num_surfels, int32_t width);

/* Per candidate: conj(column) . residual and |column|^2, without area; out [S, 3] =
 * (corr.re, corr.im, norm). residual [A,K] complex. Deterministic (no atomics).
 * Replaces kernels/surfel.py:242-260 (correlate: surfel_correlate). */
int dsplat_surfel_correlate(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                            const float* centers, const float* normals, int32_t num_surfels,
                            const float* residual, float* out, void* workspace,
                            size_t workspace_bytes, void* stream);

/* Correlation, norm and their center (and normal) derivatives: out [S, 3 + 3w] =
 * (corr.re, corr.im, norm, dcorr.re[w], dcorr.im[w], dnorm[w]), w = 6 with
 * include_normals else 3. Replaces kernels/surfel.py:293-315
 * (correlate_gradient: surfel_correlate_gradient). */
int dsplat_surfel_correlate_gradient(const dsplat_fmcw* cfg, const float* positions,
                                     int32_t num_apertures, const float* centers,
                                     const float* normals, int32_t num_surfels,
                                     const float* residual, int32_t include_normals, float* out,
                                     void* workspace, size_t workspace_bytes, void* stream);

/* Correlations and norms for each surfel's own normal and B bank normals: out
 * [S, 1 + B, 3]. With range_table [A, n_fft*oversample] complex and norm_table
 * [n_fft*oversample], the bin sum is a 4-tap Catmull-Rom lookup (residual unused).
 * Replaces kernels/surfel.py:263-290 (correlate_orientations). */
int dsplat_surfel_correlate_orientations(const dsplat_fmcw* cfg, const float* positions,
                                         int32_t num_apertures, const float* centers,
                                         const float* normals, int32_t num_surfels,
                                         const float* bank, int32_t num_normals,
                                         const float* residual, const float* range_table,
                                         const float* norm_table, int32_t oversample, float* out,
                                         void* workspace, size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DSPLAT_HIP_H */
