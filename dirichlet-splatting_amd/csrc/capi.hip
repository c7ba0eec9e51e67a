// C ABI of libdsplat_hip.so: argument validation and dispatch to the launch units.
#include <math.h>

#include <mutex>

#include "dsplat_internal.h"

namespace dsplat {

// Scratch of dsplat_surfel_vjp_noalloc: file-scope device memory that every device holds once
// its code object is loaded, until the process ends. In host code `scratch` names a host-side
// stand-in; the device address of each device's copy comes from hipGetSymbolAddress (taken on
// the first call on that device, with that device current) and is the only form host code
// passes on.
static __device__ __align__(256) unsigned char scratch[DSPLAT_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_lock;
static void* g_scratch_address[kMaxDevices];

// Device address of the current device's scratch.
static hipError_t scratch_address(void** out) {
    int device = -1;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return e;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> guard(g_scratch_lock);
    if (!g_scratch_address[device]) {
        void* address = nullptr;
        e = hipGetSymbolAddress(&address, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!address) return hipErrorInvalidSymbol;
        g_scratch_address[device] = address;
    }
    *out = g_scratch_address[device];
    return hipSuccess;
}

// Zero fill without hipMemsetAsync, so a caller may capture these calls in a HIP graph.
__global__ void zero_fill(float* __restrict__ p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

hipError_t fill_zero(float* p, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(zero_fill, dim3((unsigned)blocks), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t allow_lds(const void* kernel, size_t bytes) {
    if (bytes <= 65536) return hipSuccess;
    if (bytes > 160 * 1024) return hipErrorInvalidValue;
    return hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

bool make_chirp(const dsplat_fmcw* cfg, Chirp* out) {
    if (!cfg) return false;
    const int S = cfg->samples, P = cfg->pad, N = cfg->n_fft, K = cfg->num_bins;
    if (S < 1 || !(P == 1 || P == 2 || P == 4 || P == 8) || N != S * P || K != N / 2) return false;
    if (N < 32 || K > 512) return false;
    if (cfg->splat_form != 0 && cfg->splat_form != 1) return false;
    if (!(cfg->k0_per_m > 0) || !(cfg->c0_m_s > 0) || !(cfg->lambda_m > 0)) return false;
    Chirp c;
    c.K = K;
    c.N = N;
    c.S = S;
    c.P = P;
    c.form = cfg->splat_form;
    c.Kp = (K + 15) / 16 * 16;
    c.k0pm = cfg->k0_per_m;
    c.fc = cfg->fc_hz;
    c.slope = cfg->slope_hz_s;
    c.t0 = cfg->t_start_s;
    c.c0 = cfg->c0_m_s;
    c.two_over_c0 = 2.0 / cfg->c0_m_s;
    c.inv_N = 1.0 / N;
    c.two_pi_N = (float)(2.0 * M_PI / N);
    c.lam_d = cfg->lambda_m;
    c.fresnel_d = cfg->fresnel_n;
    c.lam = (float)cfg->lambda_m;
    c.fresnel_n = (float)cfg->fresnel_n;
    c.inv_S = (float)(1.0 / S);
    c.pi_P = (float)(M_PI / P);
    c.pi_N = (float)(M_PI / N);
    c.h = (float)(M_PI * (S - 1) / N);
    c.k0pm_f = (float)cfg->k0_per_m;
    // carrier rate 2 pi 2 / c0 (fc + K (t0 - 2 r / c0)) = rate0 - rate1 r (scattering.slang:47-50)
    c.rate0 = (float)(4.0 * M_PI / cfg->c0_m_s * (cfg->fc_hz + cfg->slope_hz_s * cfg->t_start_s));
    c.rate1 = (float)(4.0 * M_PI / cfg->c0_m_s * cfg->slope_hz_s * 2.0 / cfg->c0_m_s);
    const double a = M_PI / P, b = M_PI / N;
    c.c2 = (float)((b * b - a * a) / 6.0);
    c.c4 = (float)((3 * a * a * a * a + 7 * b * b * b * b - 10 * a * a * b * b) / 360.0);
    *out = c;
    return true;
}

}  // namespace dsplat

using namespace dsplat;

extern "C" {

const char* dsplat_version(void) { return "dsplat-hip 0.2.0 gfx950"; }

const char* dsplat_error_string(int code) { return hipGetErrorString((hipError_t)code); }

int dsplat_surfel_render(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                         const float* centers, const float* normals, const float* weights, int32_t num_surfels,
                         float* out, void* stream) {
    Chirp c;
    if (!make_chirp(cfg, &c) || num_apertures < 0 || num_surfels < 0) return hipErrorInvalidValue;
    if (num_apertures == 0) return hipSuccess;
    if (!positions || !out || (num_surfels > 0 && (!centers || !normals || !weights))) return hipErrorInvalidValue;
    if (num_surfels == 0) {
        return fill_zero(out, 2 * (size_t)num_apertures * c.K, (hipStream_t)stream);
    }
    return launch_render(c, positions, num_apertures, centers, normals, reinterpret_cast<const float2*>(weights),
                         num_surfels, reinterpret_cast<float2*>(out), (hipStream_t)stream);
}

int dsplat_surfel_jvp(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures, const float* centers,
                      const float* normals, const float* weights, const float* delta_centers,
                      const float* delta_normals, const float* delta_weights, int32_t num_surfels, float* out,
                      void* stream) {
    Chirp c;
    if (!make_chirp(cfg, &c) || num_apertures < 0 || num_surfels < 0) return hipErrorInvalidValue;
    if (num_apertures == 0) return hipSuccess;
    if (!positions || !out) return hipErrorInvalidValue;
    if (num_surfels == 0) return fill_zero(out, 2 * (size_t)num_apertures * c.K, (hipStream_t)stream);
    if (!centers || !normals || !weights || !delta_centers || !delta_normals || !delta_weights)
        return hipErrorInvalidValue;
    return launch_jvp(c, positions, num_apertures, centers, normals, reinterpret_cast<const float2*>(weights),
                      delta_centers, delta_normals, reinterpret_cast<const float2*>(delta_weights), num_surfels,
                      reinterpret_cast<float2*>(out), (hipStream_t)stream);
}

size_t dsplat_surfel_vjp_workspace(const dsplat_fmcw* cfg, int32_t num_apertures, int32_t num_surfels) {
    Chirp c;
    if (!make_chirp(cfg, &c) || num_apertures <= 0 || num_surfels <= 0) return 0;
    if (use_td(c)) return td_vjp_workspace(c, num_apertures, num_surfels);
    const VjpShape v = vjp_shape(num_apertures, num_surfels);
    return sizeof(float) * 8 * (size_t)num_surfels * v.nblk;
}

int dsplat_surfel_vjp(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures, const float* centers,
                      const float* normals, const float* weights, int32_t num_surfels, const float* upstream,
                      float* grad_centers, float* grad_normals, float* grad_weights, void* workspace,
                      size_t workspace_bytes, void* stream) {
    Chirp c;
    if (!make_chirp(cfg, &c) || num_apertures < 0 || num_surfels < 0) return hipErrorInvalidValue;
    if (num_surfels == 0) return hipSuccess;
    hipStream_t st = (hipStream_t)stream;
    if (!grad_centers || !grad_normals || !grad_weights) return hipErrorInvalidValue;
    if (num_apertures == 0) {
        hipError_t e = fill_zero(grad_centers, 3 * (size_t)num_surfels, st);
        if (e == hipSuccess) e = fill_zero(grad_normals, 3 * (size_t)num_surfels, st);
        if (e == hipSuccess) e = fill_zero(grad_weights, 2 * (size_t)num_surfels, st);
        return e;
    }
    if (!positions || !centers || !normals || !weights || !upstream || !workspace) return hipErrorInvalidValue;
    if (workspace_bytes < dsplat_surfel_vjp_workspace(cfg, num_apertures, num_surfels)) return hipErrorInvalidValue;
    return launch_vjp(c, positions, num_apertures, centers, normals, reinterpret_cast<const float2*>(weights),
                      num_surfels, reinterpret_cast<const float2*>(upstream), static_cast<float*>(workspace),
                      grad_centers, grad_normals, grad_weights, st);
}

size_t dsplat_noalloc_scratch_bytes(void) { return DSPLAT_NOALLOC_SCRATCH_BYTES; }

int dsplat_surfel_vjp_noalloc(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                              const float* centers, const float* normals, const float* weights, int32_t num_surfels,
                              const float* upstream, float* grad_centers, float* grad_normals, float* grad_weights,
                              void* stream) {
    Chirp c;
    if (!make_chirp(cfg, &c) || num_apertures < 0 || num_surfels < 0) return hipErrorInvalidValue;
    if (num_surfels == 0) return hipSuccess;
    if (!grad_centers || !grad_normals || !grad_weights) return hipErrorInvalidValue;
    if (num_apertures == 0)
        return dsplat_surfel_vjp(cfg, positions, 0, centers, normals, weights, num_surfels, upstream, grad_centers,
                                 grad_normals, grad_weights, nullptr, 0, stream);
    const size_t need = dsplat_surfel_vjp_workspace(cfg, num_apertures, num_surfels);
    if (need == 0) return hipErrorInvalidValue;
    if (need > DSPLAT_NOALLOC_SCRATCH_BYTES) return hipErrorOutOfMemory;
    void* address = nullptr;
    hipError_t e = scratch_address(&address);
    if (e != hipSuccess) return e;
    // The VJP kernels write every partial, operand and scale they read back (nothing is
    // accumulated into the workspace), so the previous call's contents need no clearing.
    return dsplat_surfel_vjp(cfg, positions, num_apertures, centers, normals, weights, num_surfels, upstream,
                             grad_centers, grad_normals, grad_weights, address, need, stream);
}

int dsplat_surfel_design(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures, const float* centers,
                         const float* normals, const float* areas, int32_t num_surfels, int32_t real_layout,
                         float* out, void* stream) {
    Chirp c;
    if (!make_chirp(cfg, &c) || num_apertures < 0 || num_surfels < 0) return hipErrorInvalidValue;
    if (num_apertures == 0 || num_surfels == 0) return hipSuccess;
    if (!positions || !centers || !normals || !areas || !out) return hipErrorInvalidValue;
    return launch_design(c, positions, num_apertures, centers, normals, areas, num_surfels, real_layout ? 0 : 1, out,
                         (hipStream_t)stream);
}

int dsplat_surfel_jacobian(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                           const float* centers, const float* normals, int32_t num_surfels, float* columns,
                           float* derivatives, void* stream) {
    Chirp c;
    if (!make_chirp(cfg, &c) || num_apertures < 0 || num_surfels < 0) return hipErrorInvalidValue;
    if (num_apertures == 0 || num_surfels == 0) return hipSuccess;
    if (!positions || !centers || !normals || !columns || !derivatives) return hipErrorInvalidValue;
    return launch_jacobian(c, positions, num_apertures, centers, normals, num_surfels,
                           reinterpret_cast<float2*>(columns), reinterpret_cast<float2*>(derivatives),
                           (hipStream_t)stream);
}

size_t dsplat_surfel_correlate_workspace(int32_t num_apertures, int32_t num_surfels, int32_t width) {
    if (num_apertures <= 0 || num_surfels <= 0 || width <= 0) return 0;
    return sizeof(float) * (size_t)width * num_surfels * cor_shape(num_apertures, num_surfels).nblk;
}

static int correlate_common(const dsplat_fmcw* cfg, const float* positions, int32_t A, const float* centers,
                            const float* normals, int32_t S, const float* residual, int W, float* out,
                            void* workspace, size_t workspace_bytes, void* stream) {
    Chirp c;
    if (!make_chirp(cfg, &c) || A < 0 || S < 0) return hipErrorInvalidValue;
    if (S == 0) return hipSuccess;
    if (!out) return hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)stream;
    const int width = 3 + 3 * W;
    if (A == 0) return fill_zero(out, (size_t)width * S, st);
    if (!positions || !centers || !normals || !residual || !workspace) return hipErrorInvalidValue;
    if (workspace_bytes < dsplat_surfel_correlate_workspace(A, S, width)) return hipErrorInvalidValue;
    return launch_correlate(c, W, positions, A, centers, normals, S, reinterpret_cast<const float2*>(residual),
                            static_cast<float*>(workspace), out, st);
}

int dsplat_surfel_correlate(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                            const float* centers, const float* normals, int32_t num_surfels, const float* residual,
                            float* out, void* workspace, size_t workspace_bytes, void* stream) {
    return correlate_common(cfg, positions, num_apertures, centers, normals, num_surfels, residual, 0, out, workspace,
                            workspace_bytes, stream);
}

int dsplat_surfel_correlate_gradient(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                                     const float* centers, const float* normals, int32_t num_surfels,
                                     const float* residual, int32_t include_normals, float* out, void* workspace,
                                     size_t workspace_bytes, void* stream) {
    return correlate_common(cfg, positions, num_apertures, centers, normals, num_surfels, residual,
                            include_normals ? 6 : 3, out, workspace, workspace_bytes, stream);
}

int dsplat_surfel_correlate_orientations(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                                         const float* centers, const float* normals, int32_t num_surfels,
                                         const float* bank, int32_t num_normals, const float* residual,
                                         const float* range_table, const float* norm_table, int32_t oversample,
                                         float* out, void* workspace, size_t workspace_bytes, void* stream) {
    Chirp c;
    const int A = num_apertures, S = num_surfels, B = num_normals;
    if (!make_chirp(cfg, &c) || A < 0 || S < 0 || B < 0 || B > 255) return hipErrorInvalidValue;
    if (S == 0) return hipSuccess;
    if (!out || (B > 0 && !bank)) return hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)stream;
    const int width = 3 * (1 + B);
    if (A == 0) return fill_zero(out, (size_t)width * S, st);
    if (!positions || !centers || !normals || !workspace) return hipErrorInvalidValue;
    if (range_table ? (!norm_table || oversample < 1) : !residual) return hipErrorInvalidValue;
    if (workspace_bytes < dsplat_surfel_correlate_workspace(A, S, width)) return hipErrorInvalidValue;
    return launch_orientations(c, positions, A, centers, normals, bank, B, S, reinterpret_cast<const float2*>(residual),
                               reinterpret_cast<const float2*>(range_table), norm_table, oversample,
                               static_cast<float*>(workspace), out, st);
}

int dsplat_surfel_orient_select(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                                const float* centers, const float* normals, const float* areas, int32_t num_surfels,
                                const float* bank, int32_t num_normals, const float* residual,
                                const float* range_table, const float* norm_table, int32_t oversample,
                                int32_t real_part, double* best_value, int32_t* best_index, void* stream) {
    Chirp c;
    const int A = num_apertures, S = num_surfels, B = num_normals;
    if (!make_chirp(cfg, &c) || A < 1 || S < 0 || B < 0 || B > 24) return hipErrorInvalidValue;
    if (S == 0) return hipSuccess;
    if (!best_value || !best_index || !areas || (B > 0 && !bank)) return hipErrorInvalidValue;
    if (!positions || !centers || !normals) return hipErrorInvalidValue;
    if (range_table ? (!norm_table || oversample < 1) : !residual) return hipErrorInvalidValue;
    const OrientPick pick = {areas, real_part != 0, best_value, reinterpret_cast<int*>(best_index)};
    return launch_orient_best(c, positions, A, centers, normals, bank, B, S, reinterpret_cast<const float2*>(residual),
                              reinterpret_cast<const float2*>(range_table), norm_table, oversample, pick,
                              (hipStream_t)stream);
}

// ------------------------------------------------------------------------------ spectral
static bool make_grid(int dim, const int32_t* samples, const int32_t* nfft, int window, float radius, int taps_stride,
                      SpecGrid* g) {
    if (dim < 1 || dim > 3 || window < 0 || window > 3 || !samples || !nfft || !(radius >= 0.0f)) return false;
    g->dim = dim;
    g->window = window;
    g->radius = radius;
    g->taps_stride = taps_stride;
    for (int d = 0; d < 3; ++d) {
        g->samples[d] = d < dim ? samples[d] : 1;
        g->nfft[d] = d < dim ? nfft[d] : 1;
        if (d < dim && (samples[d] < 1 || nfft[d] < samples[d])) return false;
        if (d < dim && window == 3 && taps_stride < samples[d]) return false;
    }
    return true;
}

int dsplat_spectral_forward(int32_t dim, const int32_t* samples, const int32_t* nfft, int32_t window, float radius,
                            const float* taps, int32_t taps_stride, const float* centers, const float* coefficients,
                            int32_t num_atoms, const float* bins, int32_t num_bins, float* out, void* stream) {
    SpecGrid g;
    if (!make_grid(dim, samples, nfft, window, radius, taps_stride, &g) || num_atoms < 0 || num_bins < 0)
        return hipErrorInvalidValue;
    if (num_bins == 0) return hipSuccess;
    if (!out || !bins || (num_atoms > 0 && (!centers || !coefficients)) || (window == 3 && !taps))
        return hipErrorInvalidValue;
    return launch_spectral_forward(g, taps, centers, reinterpret_cast<const float2*>(coefficients), nullptr, nullptr,
                                   bins, num_atoms, num_bins, reinterpret_cast<float2*>(out), (hipStream_t)stream);
}

int dsplat_spectral_jvp(int32_t dim, const int32_t* samples, const int32_t* nfft, int32_t window, float radius,
                        const float* taps, int32_t taps_stride, const float* centers, const float* coefficients,
                        const float* delta_centers, const float* delta_coefficients, int32_t num_atoms,
                        const float* bins, int32_t num_bins, float* out, void* stream) {
    SpecGrid g;
    if (!make_grid(dim, samples, nfft, window, radius, taps_stride, &g) || num_atoms < 0 || num_bins < 0)
        return hipErrorInvalidValue;
    if (num_bins == 0) return hipSuccess;
    if (!out || !bins || (window == 3 && !taps)) return hipErrorInvalidValue;
    if (num_atoms > 0 && (!centers || !coefficients || !delta_centers || !delta_coefficients))
        return hipErrorInvalidValue;
    if (num_atoms == 0) return fill_zero(out, 2 * (size_t)num_bins, (hipStream_t)stream);
    return launch_spectral_forward(g, taps, centers, reinterpret_cast<const float2*>(coefficients), delta_centers,
                                   reinterpret_cast<const float2*>(delta_coefficients), bins, num_atoms, num_bins,
                                   reinterpret_cast<float2*>(out), (hipStream_t)stream);
}

int dsplat_spectral_design(int32_t dim, const int32_t* samples, const int32_t* nfft, int32_t window, float radius,
                           const float* taps, int32_t taps_stride, const float* centers, int32_t num_atoms,
                           const float* bins, int32_t num_bins, float* out, float* derivatives, void* stream) {
    SpecGrid g;
    if (!make_grid(dim, samples, nfft, window, radius, taps_stride, &g) || num_atoms < 0 || num_bins < 0)
        return hipErrorInvalidValue;
    if (num_atoms == 0 || num_bins == 0) return hipSuccess;
    if (!out || !bins || !centers || (window == 3 && !taps)) return hipErrorInvalidValue;
    return launch_spectral_design(g, taps, centers, bins, num_atoms, num_bins, reinterpret_cast<float2*>(out),
                                  reinterpret_cast<float2*>(derivatives), (hipStream_t)stream);
}

size_t dsplat_spectral_backward_workspace(int32_t dim, int32_t num_atoms, int32_t num_bins) {
    if (num_atoms <= 0 || num_bins <= 0 || dim < 1 || dim > 3) return 0;
    return sizeof(float) * (size_t)(dim + 2) * num_atoms * spectral_spans(num_bins);
}

int dsplat_spectral_backward(int32_t dim, const int32_t* samples, const int32_t* nfft, int32_t window, float radius,
                             const float* taps, int32_t taps_stride, const float* centers, const float* coefficients,
                             int32_t num_atoms, const float* bins, int32_t num_bins, const float* upstream,
                             float* out, void* workspace, size_t workspace_bytes, void* stream) {
    SpecGrid g;
    if (!make_grid(dim, samples, nfft, window, radius, taps_stride, &g) || num_atoms < 0 || num_bins < 0)
        return hipErrorInvalidValue;
    if (num_atoms == 0) return hipSuccess;
    if (!out) return hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)stream;
    if (num_bins == 0) return fill_zero(out, (size_t)(dim + 2) * num_atoms, st);
    if (!centers || !coefficients || !bins || !upstream || !workspace || (window == 3 && !taps))
        return hipErrorInvalidValue;
    if (workspace_bytes < dsplat_spectral_backward_workspace(dim, num_atoms, num_bins)) return hipErrorInvalidValue;
    return launch_spectral_backward(g, taps, centers, reinterpret_cast<const float2*>(coefficients), bins,
                                    reinterpret_cast<const float2*>(upstream), num_atoms, num_bins,
                                    static_cast<float*>(workspace), out, st);
}

// -------------------------------------------------------------------------------- search
int dsplat_search_range_gate(const float* points, int32_t count, const float* positions, int32_t num_apertures,
                             const int32_t* intervals, int32_t num_bins, float spacing, int32_t required,
                             int32_t* out, void* stream) {
    if (count < 0 || num_apertures < 0 || num_bins < 0 || !(spacing > 0.0f)) return hipErrorInvalidValue;
    if (count == 0) return hipSuccess;
    if (!points || !out || (num_apertures > 0 && (!positions || !intervals))) return hipErrorInvalidValue;
    return launch_range_gate(points, positions, intervals, count, num_apertures, num_bins, spacing, required, out,
                             (hipStream_t)stream);
}

size_t dsplat_search_peaks_workspace(int32_t count, int32_t peaks, int32_t dim) {
    if (count < 0 || peaks < 0 || dim < 1 || dim > 8) return 0;
    return peaks_workspace(count, peaks, dim);
}

int dsplat_search_peaks(const float* scores, const float* centers, int32_t count, int32_t dim, int32_t peaks,
                        float separation, int32_t* out, void* workspace, size_t workspace_bytes, void* stream) {
    if (count < 0 || peaks < 0 || dim < 1 || dim > 8) return hipErrorInvalidValue;
    if (peaks == 0) return hipSuccess;
    if (!out || !workspace || (count > 0 && (!scores || !centers))) return hipErrorInvalidValue;
    if (workspace_bytes < peaks_workspace(count, peaks, dim)) return hipErrorInvalidValue;
    return launch_peaks(scores, centers, count, peaks, dim, separation, out, workspace, workspace_bytes,
                        (hipStream_t)stream);
}

int dsplat_search_direction(const float* inverse, const float* gradient, const float* normals, const float* mask,
                            int32_t count, int32_t width, float radius, float* out, void* stream) {
    if (count < 0 || !(width == 3 || width == 6)) return hipErrorInvalidValue;
    if (count == 0) return hipSuccess;
    if (!inverse || !gradient || !mask || !out || (width == 6 && !normals)) return hipErrorInvalidValue;
    return launch_direction(inverse, gradient, normals, mask, count, width, radius, out, (hipStream_t)stream);
}

int dsplat_search_trial(const float* centers, const float* normals, const float* direction, const float* controls,
                        int32_t count, int32_t width, float fraction, float* out_centers, float* out_normals,
                        void* stream) {
    if (count < 0 || !(width == 3 || width == 6)) return hipErrorInvalidValue;
    if (count == 0) return hipSuccess;
    if (!centers || !normals || !direction || !controls || !out_centers || !out_normals) return hipErrorInvalidValue;
    return launch_trial(centers, normals, direction, controls, count, width, fraction, out_centers, out_normals,
                        (hipStream_t)stream);
}

// ------------------------------------------------------------------------- reduce, response
size_t dsplat_energy_workspace(size_t count) { return sizeof(double) * (size_t)energy_blocks_for(count); }

int dsplat_energy(const float* value, const float* reference, size_t count, float* out, void* workspace,
                  size_t workspace_bytes, void* stream) {
    if (!out || !workspace) return hipErrorInvalidValue;
    if (count > 0 && (!value || !reference)) return hipErrorInvalidValue;
    if (workspace_bytes < dsplat_energy_workspace(count)) return hipErrorInvalidValue;
    return launch_energy(reinterpret_cast<const float2*>(value), reinterpret_cast<const float2*>(reference), count,
                         static_cast<double*>(workspace), out, (hipStream_t)stream);
}

int dsplat_range_response(double* out, double low, double high, int32_t count, int32_t num_bins, int32_t samples,
                          int32_t n_fft, void* stream) {
    if (count < 0 || num_bins < 0 || samples < 1 || n_fft < samples) return hipErrorInvalidValue;
    if (count == 0 || num_bins == 0) return hipSuccess;
    if (!out) return hipErrorInvalidValue;
    return launch_range_response(reinterpret_cast<double2*>(out), low, high, count, num_bins, samples, n_fft,
                                 (hipStream_t)stream);
}

int dsplat_synthesize(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures, const float* centers,
                      const float* normals, const double* amplitudes, int32_t num_surfels, double* out,
                      void* stream) {
    Chirp c;
    if (!make_chirp(cfg, &c) || num_apertures < 0 || num_surfels < 0 || !synthesize_supported(c))
        return hipErrorInvalidValue;
    if (num_apertures == 0) return hipSuccess;
    if (!positions || !out || (num_surfels > 0 && (!centers || !normals || !amplitudes))) return hipErrorInvalidValue;
    if (num_surfels == 0) return fill_zero(reinterpret_cast<float*>(out), 4 * (size_t)num_apertures * c.K,
                                           (hipStream_t)stream);
    return launch_synthesize(c, positions, num_apertures, centers, normals,
                             reinterpret_cast<const double2*>(amplitudes), num_surfels,
                             reinterpret_cast<double2*>(out), (hipStream_t)stream);
}

// ------------------------------------------------------------------------------------ blas
int dsplat_project_design(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                          const float* centers, const float* normals, const float* areas, int32_t num_surfels,
                          const float* basis, int32_t rows, float* out, void* stream) {
    Chirp c;
    if (!make_chirp(cfg, &c) || num_apertures < 0 || num_surfels < 0 || !project_supported(c, rows))
        return hipErrorInvalidValue;
    if (num_apertures == 0 || num_surfels == 0) return hipSuccess;
    if (!positions || !centers || !normals || !areas || !basis || !out) return hipErrorInvalidValue;
    return launch_project(c, positions, num_apertures, centers, normals, areas, num_surfels, basis, rows, out,
                          (hipStream_t)stream);
}

int32_t dsplat_project_form(const dsplat_fmcw* cfg, int32_t rows) {
    Chirp c;
    if (!make_chirp(cfg, &c)) return -1;
    if (project_td_supported(c, rows)) return 0;
    return project_supported(c, rows) ? 1 : -1;
}

int dsplat_sample_basis(const dsplat_fmcw* cfg, const float* basis, int32_t rows, float* out, void* stream) {
    Chirp c;
    if (!make_chirp(cfg, &c) || !project_td_supported(c, rows)) return hipErrorInvalidValue;
    if (!basis || !out) return hipErrorInvalidValue;
    return launch_sample_basis(c, basis, rows, out, (hipStream_t)stream);
}

int dsplat_project_sampled(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                           const float* centers, const float* normals, const float* areas, int32_t num_surfels,
                           const float* sample_basis, int32_t rows, float* out, uint32_t* column_max, void* stream) {
    Chirp c;
    if (!make_chirp(cfg, &c) || num_apertures < 0 || num_surfels < 0 || !project_td_supported(c, rows))
        return hipErrorInvalidValue;
    if (num_apertures == 0 || num_surfels == 0) return hipSuccess;
    if (!positions || !centers || !normals || !areas || !sample_basis || !out) return hipErrorInvalidValue;
    return launch_project_td(c, positions, num_apertures, centers, normals, areas, num_surfels, sample_basis, rows,
                             out, column_max, (hipStream_t)stream);
}

int32_t dsplat_project_split_rows(int32_t rows) { return rows > 0 ? project_rows_pad(rows) : 0; }

int dsplat_project_sampled_split(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                                 const float* centers, const float* normals, const float* areas,
                                 int32_t num_surfels, const float* sample_basis, int32_t rows, int32_t format,
                                 const float* column_exponents, void* hi, void* lo, int64_t ld, void* stream) {
    Chirp c;
    if (!make_chirp(cfg, &c) || num_apertures < 0 || num_surfels < 0 || !project_td_supported(c, rows))
        return hipErrorInvalidValue;
    if (format != 0 && format != 1) return hipErrorInvalidValue;
    if (num_apertures == 0 || num_surfels == 0) return hipSuccess;
    if (!positions || !centers || !normals || !areas || !sample_basis || !hi || !lo) return hipErrorInvalidValue;
    if (format == 1 && !column_exponents) return hipErrorInvalidValue;
    if (ld < num_surfels) return hipErrorInvalidValue;
    if (((uintptr_t)hi | (uintptr_t)lo) % 16 != 0) return hipErrorInvalidValue;
    return launch_project_split(c, positions, num_apertures, centers, normals, areas, num_surfels, sample_basis, rows,
                                format, format == 1 ? column_exponents : nullptr, hi, lo, ld, (hipStream_t)stream);
}

size_t dsplat_syrk_split_workspace(int32_t m, int32_t n, int32_t col0) {
    if (m <= 0 || n <= 0 || col0 >= n || m % 32 != 0) return 0;
    return syrk_split_workspace(m, n, col0 < 0 ? -1 : col0);
}

int dsplat_syrk_split(int32_t format, const void* hi, const void* lo, int64_t ld, int32_t m, int32_t n, int32_t col0,
                      const float* column_exponents, float* out, void* workspace, size_t workspace_bytes,
                      void* stream) {
    if (m < 0 || n < 0 || m % 32 != 0 || col0 > n || (format != 0 && format != 1)) return hipErrorInvalidValue;
    if (m == 0 || n == 0 || col0 == n) return hipSuccess;
    if (col0 < 0) col0 = -1;
    if (!hi || !lo || !out || !workspace || ld < n) return hipErrorInvalidValue;
    if (((uintptr_t)hi | (uintptr_t)lo) % 16 != 0) return hipErrorInvalidValue;
    if (workspace_bytes < syrk_split_workspace(m, n, col0)) return hipErrorInvalidValue;
    return launch_syrk_split(format, hi, lo, ld, m, n, col0, column_exponents, out, workspace, (hipStream_t)stream);
}

size_t dsplat_qn_workspace(const dsplat_fmcw* cfg, int32_t num_apertures) {
    Chirp c;
    if (!make_chirp(cfg, &c) || num_apertures < 0) return 0;
    return qn_workspace(num_apertures, c.P, c.K) + 16;
}

int dsplat_qn_search(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures, const float* residual,
                     const float* centers, const float* normals, const float* areas, int32_t count,
                     const dsplat_qn_spec* spec, float* centers_out, float* normals_out, float* values_out,
                     void* workspace, size_t workspace_bytes, void* stream) {
    Chirp c;
    if (!make_chirp(cfg, &c) || num_apertures < 1 || count < 0 || !spec) return hipErrorInvalidValue;
    if (!(spec->width == 3 || spec->width == 6) || spec->iterations < 0 || spec->backtracks < 0)
        return hipErrorInvalidValue;
    if (count == 0) return hipSuccess;
    if (!positions || !residual || !centers || !normals || !areas || !centers_out || !normals_out || !values_out ||
        !workspace || workspace_bytes < qn_workspace(num_apertures, c.P, c.K))
        return hipErrorInvalidValue;
    return launch_qn_search(c, positions, num_apertures, reinterpret_cast<const float2*>(residual), centers, normals,
                            areas, count, *spec, centers_out, normals_out, values_out, workspace, (hipStream_t)stream);
}

size_t dsplat_qn_lockstep_workspace(const dsplat_fmcw* cfg, int32_t num_apertures, int32_t count) {
    Chirp c;
    if (!make_chirp(cfg, &c) || num_apertures < 1 || count < 1) return 0;
    return qn_lockstep_workspace(c, num_apertures, count);
}

int dsplat_qn_lockstep(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures, const void* operands,
                       const float* centers, const float* normals, const float* areas, int32_t count,
                       const dsplat_qn_spec* spec, float* centers_out, float* normals_out, float* values_out,
                       void* workspace, size_t workspace_bytes, void* stream) {
    Chirp c;
    if (!make_chirp(cfg, &c) || num_apertures < 1 || count < 0 || !spec) return hipErrorInvalidValue;
    if (!(spec->width == 3 || spec->width == 6) || spec->iterations < 0 || spec->backtracks < 0)
        return hipErrorInvalidValue;
    if (count == 0) return hipSuccess;
    const size_t need = qn_lockstep_workspace(c, num_apertures, count);
    if (need == 0 || !positions || !operands || !centers || !normals || !areas || !centers_out || !normals_out ||
        !values_out || !workspace || workspace_bytes < need)
        return hipErrorInvalidValue;
    return launch_qn_lockstep(c, positions, num_apertures, operands, centers, normals, areas, count, *spec,
                              centers_out, normals_out, values_out, workspace, (hipStream_t)stream);
}

static bool joint_layout_ok(int count, const dsplat_joint_layout* L) {
    return L && count >= 0 && L->tail >= 0 && (L->centers >= -1) && (L->normals >= -1) && (L->areas >= -1);
}

int dsplat_joint_unpack(int32_t count, const float* vector, const dsplat_joint_layout* layout,
                        const float* centers, const float* normals, const float* coefficients, const float* root,
                        const float* coefficient_scale, float* dcenters, float* dnormals, float* dweights,
                        void* stream) {
    if (!joint_layout_ok(count, layout)) return hipErrorInvalidValue;
    if (count == 0) return hipSuccess;
    if (!vector || !centers || !normals || !coefficients || !root || !coefficient_scale || !dcenters || !dnormals ||
        !dweights)
        return hipErrorInvalidValue;
    return launch_joint_unpack(count, vector, *layout, centers, normals, coefficients, root, coefficient_scale,
                               dcenters, dnormals, dweights, (hipStream_t)stream);
}

int dsplat_joint_pack(int32_t count, const float* gcenters, const float* gnormals, const float* gweights,
                      const dsplat_joint_layout* layout, const float* centers, const float* normals,
                      const float* coefficients, const float* root, const float* coefficient_scale,
                      const float* vector, const float* damping, float normalization, float* out, void* stream) {
    if (!joint_layout_ok(count, layout)) return hipErrorInvalidValue;
    if (count == 0) return hipSuccess;
    if (!gcenters || !gnormals || !gweights || !centers || !normals || !coefficients || !root ||
        !coefficient_scale || !vector || !damping || !out)
        return hipErrorInvalidValue;
    return launch_joint_pack(count, gcenters, gnormals, gweights, *layout, centers, normals, coefficients, root,
                             coefficient_scale, vector, damping, normalization, out, (hipStream_t)stream);
}

size_t dsplat_elimination_workspace(int32_t n, int32_t count) { return elimination_workspace(n, count); }

int dsplat_backward_elimination(const double* inverse, const double* b, int32_t n, int32_t count,
                                const int32_t* limited, int32_t limit, int64_t* removed, void* workspace,
                                size_t workspace_bytes, void* stream) {
    if (n < 0 || count < 0) return hipErrorInvalidValue;
    if (count == 0) return hipSuccess;
    const size_t need = elimination_workspace(n, count);
    if (need == 0 || !inverse || !b || !removed || !workspace || workspace_bytes < need) return hipErrorInvalidValue;
    return launch_elimination(inverse, b, n, count, limited, limit, reinterpret_cast<long long*>(removed), workspace,
                              (hipStream_t)stream);
}

size_t dsplat_syrk_workspace(int32_t m, int32_t n, int32_t mode) {
    if (m < 0 || n < 0 || mode < 0 || mode > 2) return 0;
    return syrk_workspace(m, n, -1, mode);
}

int dsplat_syrk(const float* rows, int32_t m, int32_t n, int32_t mode, float* out, void* workspace,
                size_t workspace_bytes, void* stream) {
    if (m < 0 || n < 0 || mode < 0 || mode > 2) return hipErrorInvalidValue;
    if (m == 0 || n == 0) return hipSuccess;
    if (!rows || !out || !workspace || workspace_bytes < syrk_workspace(m, n, -1, mode)) return hipErrorInvalidValue;
    return launch_syrk(rows, m, n, -1, mode, nullptr, out, workspace, (hipStream_t)stream);
}

int dsplat_syrk_scaled(const float* rows, int32_t m, int32_t n, int32_t mode, const uint32_t* column_max, float* out,
                       void* workspace, size_t workspace_bytes, void* stream) {
    if (m < 0 || n < 0 || mode < 0 || mode > 2) return hipErrorInvalidValue;
    if (m == 0 || n == 0) return hipSuccess;
    if (!rows || !out || !workspace || !column_max || workspace_bytes < syrk_workspace(m, n, -1, mode))
        return hipErrorInvalidValue;
    return launch_syrk(rows, m, n, -1, mode, column_max, out, workspace, (hipStream_t)stream);
}

size_t dsplat_gemm_tn_workspace(int32_t m, int32_t n, int32_t col0, int32_t mode) {
    if (m < 0 || n < 0 || col0 < 0 || col0 >= n || mode < 0 || mode > 2) return 0;
    return syrk_workspace(m, n, col0, mode);
}

int dsplat_gemm_tn(const float* rows, int32_t m, int32_t n, int32_t col0, int32_t mode, float* out, void* workspace,
                   size_t workspace_bytes, void* stream) {
    if (m < 0 || n < 0 || col0 < 0 || col0 > n || mode < 0 || mode > 2) return hipErrorInvalidValue;
    if (m == 0 || col0 == n) return hipSuccess;
    if (!rows || !out || !workspace || workspace_bytes < syrk_workspace(m, n, col0, mode))
        return hipErrorInvalidValue;
    return launch_syrk(rows, m, n, col0, mode, nullptr, out, workspace, (hipStream_t)stream);
}

int dsplat_gemm_tn_scaled(const float* rows, int32_t m, int32_t n, int32_t col0, int32_t mode,
                          const uint32_t* column_max, float* out, void* workspace, size_t workspace_bytes,
                          void* stream) {
    if (m < 0 || n < 0 || col0 < 0 || col0 > n || mode < 0 || mode > 2) return hipErrorInvalidValue;
    if (m == 0 || col0 == n) return hipSuccess;
    if (!rows || !out || !workspace || !column_max || workspace_bytes < syrk_workspace(m, n, col0, mode))
        return hipErrorInvalidValue;
    return launch_syrk(rows, m, n, col0, mode, column_max, out, workspace, (hipStream_t)stream);
}

size_t dsplat_gram_workspace(int32_t num_apertures, int32_t num_surfels) {
    if (num_apertures <= 0 || num_surfels <= 0) return 0;
    const GramShape g = gram_shape(num_apertures, num_surfels);
    return sizeof(float) * 2 * (size_t)g.splits * num_surfels * num_surfels;
}

int dsplat_gram(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures, const float* centers,
                const float* normals, const float* areas, int32_t num_surfels, int32_t mode, int32_t real_only,
                float inv_rows, const float* column_exponents, float* out, void* workspace, size_t workspace_bytes,
                void* stream) {
    Chirp c;
    const int A = num_apertures, S = num_surfels;
    if (!make_chirp(cfg, &c) || A < 0 || S < 0 || mode < 0 || mode > 2) return hipErrorInvalidValue;
    if (S == 0) return hipSuccess;
    if (!out) return hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)stream;
    if (A == 0) return fill_zero(out, (size_t)S * S * (real_only ? 1 : 2), st);
    if (!positions || !centers || !normals || !areas || !workspace) return hipErrorInvalidValue;
    if (workspace_bytes < dsplat_gram_workspace(A, S)) return hipErrorInvalidValue;
    return launch_gram(c, mode, positions, A, centers, normals, areas, mode == 2 ? column_exponents : nullptr, S,
                       real_only, inv_rows, static_cast<float*>(workspace), out, st);
}

static bool make_scanset(int num_scans, const float* const* positions, const float* const* residuals,
                         const int32_t* apertures, ScanSet* set) {
    if (num_scans < 1 || num_scans > kMaxScans || !positions || !residuals || !apertures) return false;
    set->count = num_scans;
    set->first_block[0] = 0;
    for (int i = 0; i < num_scans; ++i) {
        if (apertures[i] < 1 || !positions[i] || !residuals[i]) return false;
        set->positions[i] = positions[i];
        set->residuals[i] = reinterpret_cast<const float2*>(residuals[i]);
        set->apertures[i] = apertures[i];
        set->first_block[i + 1] = set->first_block[i] + (apertures[i] + 31) / 32;
    }
    return true;
}

size_t dsplat_scan_workspace(const dsplat_fmcw* cfg, int32_t num_scans, const int32_t* apertures,
                             int32_t num_candidates) {
    Chirp c;
    if (!make_chirp(cfg, &c) || num_scans < 1 || num_scans > 8 || !apertures || num_candidates <= 0) return 0;
    if (td_scan_supported(c)) {
        static const float dummy = 0.0f;
        const float* p[kMaxScans];
        for (int i = 0; i < kMaxScans; ++i) p[i] = &dummy;
        ScanSet set;
        if (!make_scanset(num_scans, p, p, apertures, &set)) return 0;
        return td_scan_workspace(c, set, num_candidates);
    }
    size_t blocks = 0;
    for (int i = 0; i < num_scans; ++i) blocks += (size_t)(apertures[i] + 31) / 32;
    return sizeof(float) * 3 * blocks * (size_t)num_candidates;
}

int dsplat_scan_batched(const dsplat_fmcw* cfg, int32_t num_scans, const float* const* positions,
                        const float* const* residuals, const int32_t* apertures, const float* centers,
                        const float* normals, int32_t num_candidates, int32_t chunk, float* out, void* workspace,
                        size_t workspace_bytes, void* stream) {
    Chirp c;
    ScanSet set;
    if (!make_chirp(cfg, &c) || !make_scanset(num_scans, positions, residuals, apertures, &set) || num_candidates < 0)
        return hipErrorInvalidValue;
    if (num_candidates == 0) return hipSuccess;
    if (!centers || !normals || !out || !workspace) return hipErrorInvalidValue;
    if (workspace_bytes < dsplat_scan_workspace(cfg, num_scans, apertures, num_candidates)) return hipErrorInvalidValue;
    if (td_scan_supported(c))  // chunk = apertures per range here
        return launch_td_scan(c, set, centers, normals, num_candidates, chunk, workspace, out, (hipStream_t)stream);
    if (chunk <= 0) {
        const int groups = (4096 + set.first_block[num_scans] - 1) / set.first_block[num_scans];
        chunk = (num_candidates + groups - 1) / groups;
    }
    if ((num_candidates + chunk - 1) / chunk > 65535) chunk = (num_candidates + 65534) / 65535;  // grid y limit
    chunk = (chunk + 7) / 8 * 8;
    return launch_scan(c, set, centers, normals, num_candidates, chunk, static_cast<float*>(workspace), out,
                       (hipStream_t)stream);
}

static_assert(DSPLAT_TOP_MAX == kTopMax, "the header's selection limit is the kernel's LDS capacity");

int dsplat_certificate_scores(const float* scan_out, int32_t num_scans, int32_t num_candidates, const float* areas,
                              const double* power, int32_t real_part, double* scores, void* stream) {
    if (num_scans < 1 || num_scans > 65535 || num_candidates < 0) return hipErrorInvalidValue;
    if (num_candidates == 0) return hipSuccess;
    if (!scan_out || !areas || !power || !scores) return hipErrorInvalidValue;
    return launch_certificate_scores(scan_out, num_scans, num_candidates, areas, power, real_part ? 1 : 0, scores,
                                     (hipStream_t)stream);
}

int dsplat_top_descending(const double* scores, int32_t rows, int32_t count, int32_t top, int64_t* out,
                          void* stream) {
    if (rows < 0 || count < 0 || top < 1 || top > kTopMax) return hipErrorInvalidValue;
    if (rows == 0 || count == 0) return hipSuccess;
    if (!scores || !out) return hipErrorInvalidValue;
    return launch_top_descending(scores, rows, count, top, out, (hipStream_t)stream);
}

size_t dsplat_certificate_operands_bytes(const dsplat_fmcw* cfg, int32_t num_apertures) {
    Chirp c;
    if (!make_chirp(cfg, &c) || num_apertures <= 0) return 0;
    return cert_operands_bytes(c, num_apertures);
}

int dsplat_certificate_operands(const dsplat_fmcw* cfg, const float* residual, int32_t num_apertures,
                                void* operands, size_t operands_bytes, void* stream) {
    Chirp c;
    if (!make_chirp(cfg, &c) || num_apertures <= 0 || !residual || !operands) return hipErrorInvalidValue;
    const size_t need = cert_operands_bytes(c, num_apertures);
    if (need == 0 || operands_bytes < need) return hipErrorInvalidValue;
    return launch_cert_operands(c, reinterpret_cast<const float2*>(residual), num_apertures, operands,
                                (hipStream_t)stream);
}

size_t dsplat_certificate_gradient_workspace(const dsplat_fmcw* cfg, int32_t num_apertures, int32_t num_candidates) {
    Chirp c;
    if (!make_chirp(cfg, &c)) return 0;
    return cert_grad_workspace(c, num_apertures, num_candidates);
}

int dsplat_certificate_gradient(const dsplat_fmcw* cfg, const float* positions, int32_t num_apertures,
                                const float* centers, const float* normals, int32_t num_candidates,
                                const void* operands, float* out, void* workspace, size_t workspace_bytes,
                                void* stream) {
    Chirp c;
    if (!make_chirp(cfg, &c) || num_apertures <= 0 || num_candidates < 0) return hipErrorInvalidValue;
    if (num_candidates == 0) return hipSuccess;
    if (!positions || !centers || !normals || !operands || !out || !workspace) return hipErrorInvalidValue;
    const size_t need = cert_grad_workspace(c, num_apertures, num_candidates);
    if (need == 0 || workspace_bytes < need) return hipErrorInvalidValue;
    return launch_cert_grad(c, positions, num_apertures, centers, normals, num_candidates, operands, out, workspace,
                            (hipStream_t)stream);
}

int dsplat_pcg_start(const float* rhs, const float* diagonal, int32_t n, float tolerance, float* x, float* residual,
                     float* direction, float* state, void* stream) {
    if (n < 0 || !state) return hipErrorInvalidValue;
    if (n > 0 && (!rhs || !diagonal || !x || !residual || !direction)) return hipErrorInvalidValue;
    return launch_pcg_start(rhs, diagonal, n, tolerance, x, residual, direction, state, (hipStream_t)stream);
}

int dsplat_pcg_step(const float* image, const float* diagonal, int32_t n, float* x, float* residual,
                    float* direction, float* state, void* stream) {
    if (n < 0 || !state) return hipErrorInvalidValue;
    if (n > 0 && (!image || !diagonal || !x || !residual || !direction)) return hipErrorInvalidValue;
    return launch_pcg_step(image, diagonal, n, x, residual, direction, state, (hipStream_t)stream);
}

int dsplat_cholesky(double* a, int32_t n, int32_t lda, int32_t* info, void* stream) {
    if (n < 0 || lda < n || !info) return hipErrorInvalidValue;
    if (n == 0) return fill_zero(reinterpret_cast<float*>(info), 1, (hipStream_t)stream);
    if (!a) return hipErrorInvalidValue;
    return launch_cholesky(a, n, lda, info, (hipStream_t)stream);
}

int dsplat_cholesky_f32(float* a, int32_t n, int32_t lda, int32_t* info, void* stream) {
    if (n < 0 || lda < n || !info) return hipErrorInvalidValue;
    if (n == 0) return fill_zero(reinterpret_cast<float*>(info), 1, (hipStream_t)stream);
    if (!a) return hipErrorInvalidValue;
    return launch_cholesky_f32(a, n, lda, info, (hipStream_t)stream);
}

size_t dsplat_cholesky_solve_workspace(int32_t n, int32_t nrhs) {
    if (n < 0 || nrhs < 0 || n > kCholMaxSolve) return 0;
    return cholesky_solve_workspace(n, nrhs);
}

int dsplat_cholesky_solve(const double* l, int32_t n, int32_t lda, double* b, int32_t nrhs, int32_t ldb,
                          void* workspace, size_t workspace_bytes, void* stream) {
    if (n < 0 || nrhs < 0 || lda < n || ldb < n || n > kCholMaxSolve) return hipErrorInvalidValue;
    if (n == 0 || nrhs == 0) return hipSuccess;
    if (!l || !b || !workspace || workspace_bytes < cholesky_solve_workspace(n, nrhs)) return hipErrorInvalidValue;
    return launch_cholesky_solve(l, n, lda, b, nrhs, ldb, static_cast<double*>(workspace), (hipStream_t)stream);
}

int dsplat_power_iteration(const double* h, int32_t n, int32_t ld, int32_t steps, const double* start,
                           double* out, double* work, void* stream) {
    if (n < 1 || ld < n || steps < 1 || n > kPowerMaxN || !h || !start || !out || !work)
        return hipErrorInvalidValue;
    return launch_power_iteration(h, n, ld, steps, start, out, work, (hipStream_t)stream);
}

}  // extern "C"
