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
    /* Splat form of this call (no counterpart in the reference, whose Slang kernels have one
     * form): 0 evaluates the Dirichlet kernel as its sample sum on the matrix cores where the
     * chirp has 128 or 256 ADC samples (other chirps fall back to 1), 1 uses the bin-domain
     * cotangent kernels. Both match the float64 closed form to ~1e-7. It selects the kernels,
     * and the workspace sizes, of every entry point that takes this struct. */
    int32_t splat_form;
    int32_t reserved;   /* 0 */
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

/* Workspace bytes dsplat_surfel_vjp needs for A apertures and S surfels of this chirp
 * (0 for an unsupported configuration). */
size_t dsplat_surfel_vjp_workspace(const dsplat_fmcw* cfg, int32_t num_apertures, int32_t num_surfels);

/* Gradients of Re(conj(upstream) . render) for centers [S,3], normals [S,3] and the
 * complex weights [S] (PyTorch convention: grad_weights = A^H upstream).
 * upstream [A,K] complex. Deterministic: per-block partials reduced in float64.
 * Replaces kernels/surfel.py:123-137 (vjp: vjp_bins + host double sum). */
int dsplat_surfel_vjp(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                      const float* centers, const float* normals, const float* weights,
                      int32_t num_surfels, const float* upstream, float* grad_centers,
                      float* grad_normals, float* grad_weights, void* workspace,
                      size_t workspace_bytes, void* stream);

/* dsplat_surfel_vjp without a workspace argument, for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. Its workspace is a fixed scratch array inside the library,
 * DSPLAT_NOALLOC_SCRATCH_BYTES per device, held from the library's load on a device until
 * the process ends; the call allocates nothing. Returns hipErrorOutOfMemory (2) when
 * dsplat_surfel_vjp_workspace(cfg, A, S) exceeds that size (cfg2's 128x128 apertures x 2048
 * surfels at 128 samples fit), hipErrorInvalidValue as dsplat_surfel_vjp otherwise. Every call
 * on one device shares that device's scratch: calls that overlap in time on one device (two
 * streams, two host threads) race and must be ordered by the caller. The stream must belong
 * to the current device. Same results as dsplat_surfel_vjp, bit for bit. */
#define DSPLAT_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
size_t dsplat_noalloc_scratch_bytes(void);
int dsplat_surfel_vjp_noalloc(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                              const float* centers, const float* normals, const float* weights,
                              int32_t num_surfels, const float* upstream, float* grad_centers,
                              float* grad_normals, float* grad_weights, void* stream);

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

/* The orientation choice of a candidate search in one launch: for each candidate the
 * certificate |s|^2 / max(norm, 1e-60) (float64; s = sqrt(area) a^H r, or its real part
 * with real_part; norm = area |a|^2) of its own normal and of every bank normal
 * (num_normals <= 24), and the first best: best_value [S] float64, best_index [S] int32
 * (0 = own, 1 + b = bank b). Inputs as dsplat_surfel_correlate_orientations plus areas [S];
 * one thread per candidate sums over all apertures in float32, so no workspace and no
 * [S, 1 + B] sums. Replaces the chunked orientation loop of certificate.py:73-103
 * (orient_candidates) over kernels/surfel.py:263-290. */
int dsplat_surfel_orient_select(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                                const float* centers, const float* normals, const float* areas,
                                int32_t num_surfels, const float* bank, int32_t num_normals,
                                const float* residual, const float* range_table, const float* norm_table,
                                int32_t oversample, int32_t real_part, double* best_value,
                                int32_t* best_index, void* stream);

/* Workspace bytes of dsplat_scan_batched for this chirp (0 if unsupported). */
size_t dsplat_scan_workspace(const dsplat_fmcw* cfg, int32_t num_scans, const int32_t* apertures,
                             int32_t num_candidates);

/* Certificate scan of candidates against up to 8 scans of one chirp with uneven aperture
 * counts, in one launch: out [num_scans, C, 3] = (corr.re, corr.im, norm) of every
 * candidate against each scan's residual, without area. positions[i] [A_i, 3] and
 * residuals[i] [A_i, K] complex are device pointers held in host arrays; apertures[i] =
 * A_i (host). chunk = candidates per workgroup of the bin-domain kernel, or apertures per
 * range of the sample-domain kernel used at 128 ADC samples (0: a fill heuristic; the
 * Python side autotunes it). No counterpart launch in the reference, which scores one scan at a
 * time through kernels/surfel.py:242-260 (correlate) in certificate.py:11-31. */
int dsplat_scan_batched(const dsplat_fmcw* cfg, int32_t num_scans, const float* const* positions,
                        const float* const* residuals, const int32_t* apertures, const float* centers,
                        const float* normals, int32_t num_candidates, int32_t chunk, float* out,
                        void* workspace, size_t workspace_bytes, void* stream);

/* Certificates of a batched scan with area (certificate.py:11-31): scores [num_scans, C]
 * float64 = |s|^2 area / max(norm area, 1e-60) / power[i], s = corr (real_part 0) or
 * Re corr (real_part 1), from scan_out [num_scans, C, 3] of dsplat_scan_batched, areas [C]
 * and power [num_scans] (device, float64: the residual energies, floored by the caller). */
int dsplat_certificate_scores(const float* scan_out, int32_t num_scans, int32_t num_candidates,
                              const float* areas, const double* power, int32_t real_part, double* scores,
                              void* stream);

/* Jacobi-preconditioned conjugate gradients from zero with both stopping tests on the
 * device (varpro.py:206-228 pcg, as refinement.py:193-227 calls it): the reference tests
 * curvature <= 0 and |r| <= tolerance |rhs| on the host every iteration; here a flag in
 * state makes every update after a stop a no-op, so the caller runs a fixed number of
 * iterations and nothing waits for the host (the operator and dsplat_pcg_step can be
 * captured in a HIP graph). float32 vectors [n] on the device, float64 sums.
 * dsplat_pcg_start: x = 0, residual = rhs, direction = rhs / diagonal; state [4] floats
 * (rz, running flag, goal, reserved). dsplat_pcg_step, after image = H direction: one
 * iteration's updates of x, residual, direction and state. */
int dsplat_pcg_start(const float* rhs, const float* diagonal, int32_t n, float tolerance, float* x, float* residual,
                     float* direction, float* state, void* stream);
int dsplat_pcg_step(const float* image, const float* diagonal, int32_t n, float* x, float* residual,
                    float* direction, float* state, void* stream);

/* Sample-domain certificate operands of one residual [A, K] complex: its rows taken to the
 * samples (DFT) and split into the FP16 hi / lo operand tiles of four accumulator sets per
 * aperture (the residual sums g0 and g1, the norm polynomial and its derivative), in a device
 * buffer of dsplat_certificate_operands_bytes (0: the chirp's splat form or sample count is
 * not served by the sample-domain kernels). Built once per residual and shared by every
 * dsplat_certificate_gradient launch against it; no counterpart in the reference, whose
 * correlation kernels read the residual bins directly. */
size_t dsplat_certificate_operands_bytes(const dsplat_fmcw* cfg, int32_t num_apertures);
int dsplat_certificate_operands(const dsplat_fmcw* cfg, const float* residual, int32_t num_apertures,
                                void* operands, size_t operands_bytes, void* stream);

/* Correlations, norms and their centre and normal derivatives of C candidates against the
 * residual of `operands`, in the sample domain on the FP16 matrix cores: out [C, 21] =
 * (corr.re, corr.im, norm, dcorr.re[6], dcorr.im[6], dnorm[6]), the columns of
 * dsplat_surfel_correlate_gradient with include_normals (without area). Deterministic:
 * per-range partials summed in float64. Workspace from
 * dsplat_certificate_gradient_workspace. Replaces kernels/surfel.py:293-315
 * (correlate_gradient) for chirps the sample-domain form serves. */
size_t dsplat_certificate_gradient_workspace(const dsplat_fmcw* cfg, int32_t num_apertures, int32_t num_candidates);
int dsplat_certificate_gradient(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                                const float* centers, const float* normals, int32_t num_candidates,
                                const void* operands, float* out, void* workspace, size_t workspace_bytes,
                                void* stream);

/* DSPLAT_TOP_MAX: the largest `top` of dsplat_top_descending. */
#define DSPLAT_TOP_MAX 4096

/* Indices of the min(top, count) largest entries of every row of scores [rows, count]
 * float64 in descending order, ties to the lower index, NaN first: the stable descending
 * argsort cut to its head that ranks certificates (certificate.py:52-56, 196-201), one
 * workgroup per row instead of a full sort. out [rows, min(top, count)] int64;
 * 1 <= top <= DSPLAT_TOP_MAX. */
int dsplat_top_descending(const double* scores, int32_t rows, int32_t count, int32_t top, int64_t* out,
                          void* stream);

/* ---- Spectral atoms: separable windowed Dirichlet kernels on a 1-3-D grid, bin units.
 * dim axes with samples[dim], nfft[dim]; window 0 rect, 1 hann, 2 hamming, 3 kaiser
 * (then taps [dim, taps_stride] are the normalized window taps); radius > 0 truncates.
 * centers [S, dim], coefficients [S] complex, bins [M, dim]. Gradient convention dc = -dk. */

/* out[M] complex = sum_s coeff_s a(s, m). Replaces kernels/spectral.py:168-177 (spectral_forward). */
int dsplat_spectral_forward(int32_t dim, const int32_t* samples, const int32_t* nfft, int32_t window,
                            float radius, const float* taps, int32_t taps_stride, const float* centers,
                            const float* coefficients, int32_t num_atoms, const float* bins,
                            int32_t num_bins, float* out, void* stream);

/* out[M] complex: directional derivative along delta_centers [S, dim] and
 * delta_coefficients [S] complex. Replaces kernels/spectral.py:192-201 (spectral_jvp). */
int dsplat_spectral_jvp(int32_t dim, const int32_t* samples, const int32_t* nfft, int32_t window,
                        float radius, const float* taps, int32_t taps_stride, const float* centers,
                        const float* coefficients, const float* delta_centers,
                        const float* delta_coefficients, int32_t num_atoms, const float* bins,
                        int32_t num_bins, float* out, void* stream);

/* out [M, S] complex and, when derivatives is not NULL, d/dcenters [M, S, dim] complex.
 * Replaces kernels/spectral.py:146-154 (spectral_design). */
int dsplat_spectral_design(int32_t dim, const int32_t* samples, const int32_t* nfft, int32_t window,
                           float radius, const float* taps, int32_t taps_stride, const float* centers,
                           int32_t num_atoms, const float* bins, int32_t num_bins, float* out,
                           float* derivatives, void* stream);

size_t dsplat_spectral_backward_workspace(int32_t dim, int32_t num_atoms, int32_t num_bins);

/* Gradients of Re(conj(upstream) . render): out [S, dim + 2] = (grad centers[dim],
 * grad coeff re, im). upstream [M] complex. Deterministic (no atomics).
 * Replaces kernels/spectral.py:157-165 (spectral_backward). */
int dsplat_spectral_backward(int32_t dim, const int32_t* samples, const int32_t* nfft, int32_t window,
                             float radius, const float* taps, int32_t taps_stride,
                             const float* centers, const float* coefficients, int32_t num_atoms,
                             const float* bins, int32_t num_bins, const float* upstream, float* out,
                             void* workspace, size_t workspace_bytes, void* stream);

/* ---- Spatial search. */

/* out[count] int32 = 1 where a point's round-half-even range bin lies in a marked
 * interval (intervals [A, bins] int32) at >= required apertures.
 * Replaces kernels/search.py:10-24 (candidate_range_gate). */
int dsplat_search_range_gate(const float* points, int32_t count, const float* positions,
                             int32_t num_apertures, const int32_t* intervals, int32_t num_bins,
                             float spacing, int32_t required, int32_t* out, void* stream);

/* Greedy separated peaks: out[peaks] int32 indices best first, -1 once exhausted. Same
 * selection as taking the best finite score (ties to the lower index) and dropping every
 * candidate closer than separation, repeatedly. scores [count] are not modified;
 * centers [count, dim], dim <= 8. Workspace from dsplat_search_peaks_workspace.
 * Replaces kernels/search.py:27-39 (candidate_peaks). */
size_t dsplat_search_peaks_workspace(int32_t count, int32_t peaks, int32_t dim);
int dsplat_search_peaks(const float* scores, const float* centers, int32_t count, int32_t dim,
                        int32_t peaks, float separation, int32_t* out, void* workspace,
                        size_t workspace_bytes, void* stream);

/* Per candidate: -H^-1 g (or -g), masked, tangent to unit normals, in the trust radius.
 * inverse [C, w, w], gradient [C, w], normals [C, 3], mask [w], w = 3 or 6.
 * Replaces kernels/search.py:42-56 (candidate_direction). */
int dsplat_search_direction(const float* inverse, const float* gradient, const float* normals,
                            const float* mask, int32_t count, int32_t width, float radius, float* out,
                            void* stream);

/* Trial step: centers clamped to the box where trainable, normals renormalized.
 * controls [4, 6]: scale, trainable mask, lower, upper per coordinate.
 * Replaces kernels/search.py:59-77 (candidate_trial). */
int dsplat_search_trial(const float* centers, const float* normals, const float* direction,
                        const float* controls, int32_t count, int32_t width, float fraction,
                        float* out_centers, float* out_normals, void* stream);

/* ---- Reductions and responses. */

size_t dsplat_energy_workspace(size_t count);

/* out[0] = sum |value - reference|^2 over count complex entries, block sums added in
 * float64 in a fixed order. Replaces kernels/reduce.py:8-18 (complex_energy + sum). */
int dsplat_energy(const float* value, const float* reference, size_t count, float* out,
                  void* workspace, size_t workspace_bytes, void* stream);

/* Float64 FMCW scan of surfels with complex amplitudes by the ADC signal chain, not the
 * Dirichlet kernel: each aperture-surfel echo gain e^{i carrier} b_s (cfg->fresnel_n > 1 adds
 * the Fresnel factor) is a beat tone of 2 slope r / c0 sampled at the ADC rate; the tones are
 * summed, zero-padded to n_fft, transformed, cut to the positive bins and divided by
 * samples. amplitudes [S] complex128 (interleaved doubles); out [A, K] complex128.
 * samples at most 256. Replaces groundtruth.py:58-90
 * (synthesize: a torch float64 loop over aperture and surfel blocks). */
int dsplat_synthesize(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                      const float* centers, const float* normals, const double* amplitudes,
                      int32_t num_surfels, double* out, void* stream);

/* out [bins, count] complex128: float64 D(bin - center) for count centers spread over
 * [low, high]. Replaces kernels/response.py:8-20 (range_response). */
int dsplat_range_response(double* out, double low, double high, int32_t count, int32_t num_bins,
                          int32_t samples, int32_t n_fft, void* stream);

/* Row-projected real dictionary block in one pass: out [rows * A, S] row-major, row
 * j * A + a = sum over (k, part) of basis [rows, 2K][j][2k + part] times the real (part 0)
 * or imaginary (part 1) part of atom (a, k, s) = sqrt(area_s) amp rot D(k - k0). basis is
 * the interleaved real form of a row basis U^H (rows even); atoms are never written.
 * Replaces the design + reduce_major product of the streamed blocks (fmcw.py:243-261,
 * surfel.slang:313-378). */
int dsplat_project_design(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                          const float* centers, const float* normals, const float* areas,
                          int32_t num_surfels, const float* basis, int32_t rows, float* out,
                          void* stream);

/* Which projection kernel serves cfg and rows: 0 for dsplat_project_sampled (cfg->splat_form
 * 0 and 128 or 256 ADC samples), 1 for dsplat_project_design only, -1 for
 * neither (invalid configuration or odd rows). */
int32_t dsplat_project_form(const dsplat_fmcw* cfg, int32_t rows);

/* Sample-domain form of a row basis: out [rows, 2 * samples] = basis [rows, 2K] times the real
 * 2x2-block form of F[k][n] = e^{-2 pi i n k / N}, summed in float64 and rounded once; row
 * pairs keep the interleaved (Re, Im) layout. The input of dsplat_project_sampled; a basis is
 * converted once and reused for every block. Same role as dsplat_project_design's basis
 * (fmcw.py:95-98 reduce_major). */
int dsplat_sample_basis(const dsplat_fmcw* cfg, const float* basis, int32_t rows, float* out, void* stream);

/* dsplat_project_design's output (same layout, same rows) from the sample-domain basis of
 * dsplat_sample_basis: the atoms' sample-domain powers v^n are projected on the matrix cores
 * and never formed per bin. column_max (NULL or [num_surfels], zeroed by the caller) receives
 * max(column_max[s], max |out[., s]|) as float bits: the column maxima that FP16 mode of
 * dsplat_syrk_scaled / dsplat_gemm_tn_scaled takes, without another pass over out.
 * hipErrorInvalidValue unless dsplat_project_form is 0. Replaces the design + reduce_major
 * product of the streamed blocks (fmcw.py:243-261, surfel.slang:313-378). */
int dsplat_project_sampled(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                           const float* centers, const float* normals, const float* areas,
                           int32_t num_surfels, const float* sample_basis, int32_t rows, float* out,
                           uint32_t* column_max, void* stream);

/* Rows per aperture of a split projected block: rows rounded up to 32. */
int32_t dsplat_project_split_rows(int32_t rows);

/* dsplat_project_sampled's block in the operand form of dsplat_syrk_split: 16-bit hi and lo
 * arrays (16-byte aligned) with hi + lo ~ the float32 value (hi = RNE(value), lo =
 * RNE(value - hi)), BF16 for format 0, FP16 of the value times 2^column_exponents[s] for
 * format 1 (float, integral values [S]: they keep the column inside the FP16 range), in
 * 32-row column tiles: element (k, s) at ((k / 32) ld + s) 32 + k % 32, ld >= S,
 * m = A * rows_pad rows (rows_pad = dsplat_project_split_rows(rows)); projected row j of
 * aperture a is row k = a * rows_pad + j, and rows j >= rows are written as zeros. Replaces the design + reduce_major product
 * of the streamed blocks (fmcw.py:243-261, surfel.slang:313-378) for the split SYRK. */
int dsplat_project_sampled_split(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                                 const float* centers, const float* normals, const float* areas,
                                 int32_t num_surfels, const float* sample_basis, int32_t rows, int32_t format,
                                 const float* column_exponents, void* hi, void* lo, int64_t ld, void* stream);

/* Upper triangle of out [n, n] += P^T P (col0 < 0), or out [n, n - col0] += P^T P[:, col0:]
 * (col0 >= 0: the cross products of the trailing columns with all columns), for P [m, n]
 * given as 16-bit hi / lo arrays in 32-row column tiles (element (k, c) at
 * ((k / 32) ld + c) 32 + k % 32, ld >= n, m a multiple of 32, 16-byte aligned): BF16 for
 * format 0; FP16 of column c scaled by 2^column_exponents[c] for format 1 (the products are
 * unscaled by 2^-(e_i + e_j)). Products hi hi + hi lo + lo hi on MFMA with FP32 accumulation
 * over at most 32768 rows, split-K partials summed in float64 in a fixed order. Deterministic. Workspace from
 * dsplat_syrk_split_workspace. Replaces kernels/cublas.py:31-44 (cublasSsyrk_v2) and the
 * streamed addmm_ of replacement.py:178-199 on projected blocks. */
size_t dsplat_syrk_split_workspace(int32_t m, int32_t n, int32_t col0);
int dsplat_syrk_split(int32_t format, const void* hi, const void* lo, int64_t ld, int32_t m, int32_t n, int32_t col0,
                      const float* column_exponents, float* out, void* workspace, size_t workspace_bytes,
                      void* stream);

/* ---- Batched quasi-Newton certificate search. */

/* Per-search constants of dsplat_qn_search, from the parameter specs (certificate.py:75-86
 * _controls and parameters.py:57-70 tangent). Lower/upper are +-inf when unbounded. */
typedef struct dsplat_qn_spec {
    float center_scale, normal_scale;
    float center_mask[3], normal_mask[3]; /* 1 = trainable */
    float direction_mask[6];              /* controls[1] */
    float step_scale[6];                  /* controls[0] */
    float lower[3], upper[3];             /* controls[2], controls[3] */
    int32_t width;                        /* 3: centers; 6: centers and unit normals */
    int32_t normalize;                    /* the normal field is a unit vector */
    int32_t real_part;                    /* certificate of Re(a^H r) */
    int32_t iterations, backtracks;
    float radius;                         /* trust radius of a direction */
    double power;                         /* |r|^2 (clamped) */
} dsplat_qn_spec;

/* Quasi-Newton ascent of the residual certificate for `count` candidates in one launch:
 * from centers/normals [count, 3], each candidate takes damped BFGS directions with halving
 * backtracks until its certificate stops rising or `iterations` steps are done. Outputs
 * the final centers, normals [count, 3] and certificate values [count] (float32). residual
 * [A, K] complex64 interleaved; areas [count]. Workspace from dsplat_qn_workspace.
 * Replaces the host loop of certificate.py:106-201 (search_certificate). */
size_t dsplat_qn_workspace(const dsplat_fmcw* cfg, int32_t num_apertures);
int dsplat_qn_search(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                     const float* residual, const float* centers, const float* normals, const float* areas,
                     int32_t count, const dsplat_qn_spec* spec, float* centers_out, float* normals_out,
                     float* values_out, void* workspace, size_t workspace_bytes, void* stream);

/* The same ascent, with the same outputs and stopping rules, on chirps the sample-domain
 * kernels serve, in lockstep rounds: each round evaluates the certificate and its gradient of
 * every live candidate at its next point in one dsplat_certificate_gradient launch against
 * `operands` (dsplat_certificate_operands of the residual), and one thread per candidate
 * accepts the point (with the BFGS update and the next direction) or halves the step. An
 * iteration whose first trial rises costs one round. The call reads a live count from the
 * device every few rounds, so it synchronizes `stream` and cannot be graph-captured.
 * Workspace from dsplat_qn_lockstep_workspace (0: chirp not served). Replaces the host loop
 * of certificate.py:106-201 (search_certificate), like dsplat_qn_search. */
size_t dsplat_qn_lockstep_workspace(const dsplat_fmcw* cfg, int32_t num_apertures, int32_t count);
int dsplat_qn_lockstep(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures, const void* operands,
                       const float* centers, const float* normals, const float* areas, int32_t count,
                       const dsplat_qn_spec* spec, float* centers_out, float* normals_out, float* values_out,
                       void* workspace, size_t workspace_bytes, void* stream);

/* ---- Dense products. */

/* Upper triangle of out [n, n] (row-major) += rows^T rows, rows [m, n] row-major float32,
 * on MFMA; the lower triangle is untouched. mode 0: FP32 operands; 1: BF16 hi + lo split
 * (3 products); 2: FP16 hi + lo split with per-column power-of-two scaling (3 products);
 * FP32 accumulation, float64 split-K reduction. Deterministic. Workspace from
 * dsplat_syrk_workspace. Replaces kernels/cublas.py:31-44 (syrk: cublasSsyrk_v2 through
 * ctypes). */
/* Float64 Cholesky factor of the K x K systems, in place: on return the lower triangle of a
 * (row-major, leading dimension lda) holds L with A = L L^T and the strictly upper part is
 * zero; only the lower triangle is read. *info (device int) = 1-based index of the first
 * non-positive pivot, 0 on success; after a failure the factor is not meaningful (the later
 * block columns are skipped). Replaces torch.linalg.cholesky_ex (cuSOLVER potrf) of
 * varpro.py:44-79 (solve, shifted_cholesky). */
int dsplat_cholesky(double* a, int32_t n, int32_t lda, int32_t* info, void* stream);

/* The same factorization in float32: the first attempt of varpro.py:62-79 (solve), which
 * the reference runs as torch.linalg.cholesky_ex (cuSOLVER potrf) on the float32 Gram.
 * info as for dsplat_cholesky, from the float32 pivots. */
int dsplat_cholesky_f32(float* a, int32_t n, int32_t lda, int32_t* info, void* stream);

/* Solve L L^T X = B in place for nrhs right-hand sides, column j of B at b + j * ldb, with
 * L from dsplat_cholesky: forward and backward substitution by 64-row blocks, one launch per
 * block and direction; the workspace (dsplat_cholesky_solve_workspace bytes) holds the forward
 * solution. Replaces torch.cholesky_solve (cuSOLVER potrs) of varpro.py:62-79. */
size_t dsplat_cholesky_solve_workspace(int32_t n, int32_t nrhs);
int dsplat_cholesky_solve(const double* l, int32_t n, int32_t lda, double* b, int32_t nrhs, int32_t ldb,
                          void* workspace, size_t workspace_bytes, void* stream);

/* Power iteration of the largest-eigenvalue estimate: `steps` products y_i = h x_i of the
 * float64 [n, n] matrix h (row-major, leading dimension ld), x_0 = start, x_i = y_(i-1) /
 * max(|y_(i-1)|, DBL_MIN) (the tiny of float64); out [n] = y_(steps-1), whose norm is the
 * estimate. work holds 2 n doubles; n <= 16384. One launch per step: every workgroup
 * normalizes the previous product itself. Replaces the host loop of varpro.py:31-41
 * (largest_eigenvalue) for real float64 systems. */
int dsplat_power_iteration(const double* h, int32_t n, int32_t ld, int32_t steps, const double* start,
                           double* out, double* work, void* stream);

/* ---- Joint Gauss-Newton operator: the parameter maps around the splat's JVP and VJP. */

/* Layout of the dimensionless unknowns of the joint system (refinement.py:113-121): the
 * first unknown of each trainable field's section (count x 3 for centers and normals, count
 * for areas; -1 when the field is not trainable; every component of a field is free), then
 * the coefficients from `tail` (one real value per atom, or (Re, Im) pairs). */
typedef struct dsplat_joint_layout {
    int32_t centers, normals, areas, tail;
    int32_t complex_mode;               /* 1: (Re, Im) pair per coefficient */
    int32_t centers_unit, normals_unit; /* the field holds unit vectors (tangent projection) */
    float center_scale, normal_scale, area_scale; /* physical length of one solver unit */
} dsplat_joint_layout;

/* Unknowns -> splat tangents (refinement.py:136-146 unpack, surfel.py:53-61 weight chain):
 * dcenters, dnormals [count, 3] float32 (zero for an untrained field), dweights [count] complex
 * (interleaved float32) = db sqrt(area) + b d(area) / (2 sqrt(area)). coefficients [count]
 * complex interleaved, root = sqrt(area), coefficient_scale [count]. */
int dsplat_joint_unpack(int32_t count, const float* vector, const dsplat_joint_layout* layout,
                        const float* centers, const float* normals, const float* coefficients, const float* root,
                        const float* coefficient_scale, float* dcenters, float* dnormals, float* dweights,
                        void* stream);

/* The VJP's gradients (gcenters, gnormals [count, 3], gweights [count] complex interleaved) ->
 * the operator's image (refinement.py:148-158 adjoint, 170-174 matvec): out = adjoint /
 * normalization + damping * vector over every unknown; damping = ridge diagonal + damping. */
int dsplat_joint_pack(int32_t count, const float* gcenters, const float* gnormals, const float* gweights,
                      const dsplat_joint_layout* layout, const float* centers, const float* normals,
                      const float* coefficients, const float* root, const float* coefficient_scale,
                      const float* vector, const float* damping, float normalization, float* out, void* stream);

/* Backward elimination of `count` unknowns from a least-squares solution, one launch per
 * step on the stream (no host synchronization): inverse [n, n] float64 (of the regularized
 * real Gram), b [n] = inverse q. Each step removes the unknown with the smallest
 * b_j^2 / W_jj (ties to the lower index) among those still allowed, keeping the exact
 * inverse of the rest by rank-one corrections; at most `limit` unknowns with
 * limited[j] != 0 leave (limited may be NULL). removed [count] int64 in order.
 * count <= n, count <= 4096. Replaces the host loop of replacement.py:104-139
 * (backward_elimination). */
size_t dsplat_elimination_workspace(int32_t n, int32_t count);
int dsplat_backward_elimination(const double* inverse, const double* b, int32_t n, int32_t count,
                                const int32_t* limited, int32_t limit, int64_t* removed, void* workspace,
                                size_t workspace_bytes, void* stream);

size_t dsplat_syrk_workspace(int32_t m, int32_t n, int32_t mode);
int dsplat_syrk(const float* rows, int32_t m, int32_t n, int32_t mode, float* out, void* workspace,
                size_t workspace_bytes, void* stream);
/* dsplat_syrk with the column maxima of rows given (float bits of an upper bound of
 * max |rows[., c]| per column, as dsplat_project_sampled writes them): FP16 mode (2) then
 * skips its own pass over rows; other modes ignore them. Same workspace. Replaces the same
 * cuBLAS call (kernels/cublas.py:31-44) on projected blocks. */
int dsplat_syrk_scaled(const float* rows, int32_t m, int32_t n, int32_t mode, const uint32_t* column_max,
                       float* out, void* workspace, size_t workspace_bytes, void* stream);

/* out [n, n - col0] (row-major) += rows^T rows[:, col0:], rows [m, n] float32: the cross
 * products of the trailing columns with all columns, on the SYRK kernel's tiles and modes.
 * Replaces the streamed addmm_ of the pooled statistics (replacement.py:178-199). */
size_t dsplat_gemm_tn_workspace(int32_t m, int32_t n, int32_t col0, int32_t mode);
int dsplat_gemm_tn(const float* rows, int32_t m, int32_t n, int32_t col0, int32_t mode, float* out,
                   void* workspace, size_t workspace_bytes, void* stream);
/* dsplat_gemm_tn with the column maxima of rows given, as dsplat_syrk_scaled. Replaces the
 * same streamed addmm_ (replacement.py:178-199) on projected blocks. */
int dsplat_gemm_tn_scaled(const float* rows, int32_t m, int32_t n, int32_t col0, int32_t mode,
                          const uint32_t* column_max, float* out, void* workspace, size_t workspace_bytes,
                          void* stream);

size_t dsplat_gram_workspace(int32_t num_apertures, int32_t num_surfels);

/* Gram of the surfel dictionary, atoms (with sqrt(area)) generated on chip and contracted
 * on MFMA: out = inv_rows * A^H A, complex [S, S] Hermitian, or its real part [S, S]
 * float when real_only. Rows are every aperture's K bins. mode 0 FP32 MFMA; 1 BF16
 * hi+lo split (3 products, FP32 accumulation); 2 FP16 hi+lo split with per-column
 * power-of-two scales 2^column_exponents[s] (float, integral values; NULL = 1).
 * Split-K partials are summed in float64 in a fixed order. Replaces the streamed
 * normal equations of varpro.py:91-120 (design blocks + cublasSsyrk) for the Gram. */
int dsplat_gram(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                const float* centers, const float* normals, const float* areas, int32_t num_surfels,
                int32_t mode, int32_t real_only, float inv_rows, const float* column_exponents,
                float* out, void* workspace, size_t workspace_bytes, void* stream);

/* ---- Measurement aid (no counterpart in the reference). */

/* VALU issue-rate calibration: blocks x 256 lanes each run iters steps of eight independent
 * FMA chains in FP32 (fp64 = 0) or FP64 (fp64 = 1), 8 x iters wave64 FMA instructions per
 * wave; out receives one float (or double) per lane, blocks x 256. For bench.py's valu workload. */
int dsplat_calibrate_valu(int32_t fp64, int32_t blocks, int32_t iters, void* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DSPLAT_HIP_H */
