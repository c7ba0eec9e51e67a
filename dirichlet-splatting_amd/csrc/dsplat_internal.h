// Host-side declarations shared by the launch units and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/dsplat_hip.h"
#include "dsplat_device.h"

namespace dsplat {

// Validated kernel constants from the ABI struct; false if the configuration is unsupported.
bool make_chirp(const dsplat_fmcw* cfg, Chirp* out);
hipError_t fill_zero(float* p, size_t n, hipStream_t st);
// Raise a kernel's dynamic LDS limit above 64 KB (gfx950 has 160 KB per workgroup).
hipError_t allow_lds(const void* kernel, size_t bytes);

struct VjpShape {
    int nblk;    // aperture blocks of 32
    int groups;  // surfel chunks
    int chunk;   // surfels per chunk
};
VjpShape vjp_shape(int A, int S);

struct CorShape {
    int nblk;    // aperture blocks of 32
    int groups;  // candidate chunks
    int chunk;   // candidates per chunk
};
// Aperture blocks and candidate chunks for about `target` workgroups (the workspace depends on
// the blocks only). Correlations aim at one round of their LDS-bound workgroups.
constexpr int kCorRound = 512;
constexpr int kCorFewCandidates = 8192;
CorShape cor_shape(int A, int S, int target = 4096);

constexpr int kMaxScans = 8;
// A set of scans of one chirp with uneven aperture counts, launched together.
struct ScanSet {
    int count;
    const float* positions[kMaxScans];   // [A_i, 3]
    const float2* residuals[kMaxScans];  // [A_i, K]
    int apertures[kMaxScans];
    int first_block[kMaxScans + 1];      // prefix sums of 32-aperture blocks
};
hipError_t launch_scan(const Chirp& c, const ScanSet& set, const float* ctr, const float* nrm, int S, int chunk,
                       float* partial, float* out, hipStream_t st);
size_t cor_smem(const Chirp& c);
// Sample-domain scan (S = 128): workspace bytes and launch (chunk = apertures per range).
bool td_scan_supported(const Chirp& c);
size_t td_scan_workspace(const Chirp& c, const ScanSet& set, int C);
hipError_t launch_td_scan(const Chirp& c, const ScanSet& set, const float* ctr, const float* nrm, int C, int chunk,
                          void* ws, float* out, hipStream_t st);
// Certificate scores of a scan's (corr, norm) rows, and the top of each row of scores.
constexpr int kTopMax = 4096;  // LDS holds the selection
hipError_t launch_certificate_scores(const float* scan, int n, int C, const float* areas, const double* power,
                                     int real_part, double* scores, hipStream_t st);
hipError_t launch_top_descending(const double* scores, int rows, int count, int top, int64_t* out, hipStream_t st);
// Conjugate gradients with the stopping tests on the device (one workgroup per launch).
hipError_t launch_pcg_start(const float* rhs, const float* diagonal, int n, float tolerance, float* x, float* r,
                            float* d, float* state, hipStream_t st);
hipError_t launch_pcg_step(const float* image, const float* diagonal, int n, float* x, float* r, float* d,
                           float* state, hipStream_t st);

hipError_t launch_design(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                         const float* areas, int S, int complex_layout, float* out, hipStream_t st);
hipError_t launch_jacobian(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm, int S,
                           float2* col, float2* jac, hipStream_t st);
hipError_t launch_correlate(const Chirp& c, int W, const float* pos, int A, const float* ctr, const float* nrm, int S,
                            const float2* res, float* partial, float* out, hipStream_t st);
// Fused choice of the best orientation per candidate (instead of the [S, 1 + B, 3] sums).
struct OrientPick {
    const float* areas;
    int real_part;
    double* value;  // [S] |s|^2 / max(norm, 1e-60) of the best orientation
    int* index;     // [S] 0 = own normal, 1 + b = bank normal b
};
hipError_t launch_orientations(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                               const float* bank, int B, int S, const float2* res, const float2* table,
                               const float* norm_table, int oversample, float* partial, float* out, hipStream_t st);
hipError_t launch_orient_best(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                              const float* bank, int B, int S, const float2* res, const float2* table,
                              const float* norm_table, int oversample, const OrientPick& pick, hipStream_t st);

// Spectral grid of up to three separable axes (bin units).
struct SpecGrid {
    int dim;          // 1..3
    int window;       // 0 rect, 1 hann, 2 hamming, 3 kaiser (taps)
    int samples[3];   // observed samples per axis
    int nfft[3];      // FFT length per axis
    float radius;     // support radius in bins, 0 = exact
    int taps_stride;  // row stride of taps [dim, taps_stride]
};
constexpr int kSpectralSpan = 256;  // bins per backward partial
constexpr int kEnergyBlocks = 1024;
int spectral_spans(int M);
int energy_blocks_for(size_t n);

hipError_t launch_reduce_partials(const float* partial, int rows, int nblk, int width, float* out, hipStream_t st);
hipError_t launch_spectral_forward(const SpecGrid& g, const float* taps, const float* ctr, const float2* coeff,
                                   const float* dctr, const float2* dcoeff, const float* bins, int S, int M,
                                   float2* out, hipStream_t st);
hipError_t launch_spectral_design(const SpecGrid& g, const float* taps, const float* ctr, const float* bins, int S,
                                  int M, float2* out, float2* jac, hipStream_t st);
hipError_t launch_spectral_backward(const SpecGrid& g, const float* taps, const float* ctr, const float2* coeff,
                                    const float* bins, const float2* up, int S, int M, float* partial, float* out,
                                    hipStream_t st);
hipError_t launch_range_gate(const float* pts, const float* pos, const int* intervals, int count, int apertures,
                             int bins, float spacing, int required, int* out, hipStream_t st);
size_t peaks_workspace(int count, int peaks, int dim);
hipError_t launch_peaks(const float* scores, const float* ctr, int count, int peaks, int dim, float separation,
                        int* out, void* workspace, size_t bytes, hipStream_t st);
hipError_t launch_direction(const float* inv, const float* grad, const float* nrm, const float* mask, int count,
                            int width, float radius, float* out, hipStream_t st);
hipError_t launch_trial(const float* ctr, const float* nrm, const float* dir, const float* ctl, int count, int width,
                        float fraction, float* oc, float* on, hipStream_t st);
hipError_t launch_energy(const float2* x, const float2* y, size_t n, double* part, float* out, hipStream_t st);
bool synthesize_supported(const Chirp& c);
hipError_t launch_synthesize(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                             const double2* amp, int S, double2* out, hipStream_t st);
hipError_t launch_range_response(double2* out, double low, double high, int count, int bins, int samples, int nfft,
                                 hipStream_t st);

struct GramShape {
    int tiles;      // 128 x 128 tiles of the upper triangle
    int splits;     // aperture ranges (split-K)
    int per_split;  // apertures per range
};
GramShape gram_shape(int A, int S);
hipError_t launch_gram(const Chirp& c, int mode, const float* pos, int A, const float* ctr, const float* nrm,
                       const float* areas, const float* col_exp, int S, int real_only, float inv_rows,
                       float* partial, float* out, hipStream_t st);
bool project_supported(const Chirp& c, int rows);
hipError_t launch_project(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                          const float* areas, int S, const float* basis, int rows, float* out, hipStream_t st);
bool project_td_supported(const Chirp& c, int rows);
hipError_t launch_sample_basis(const Chirp& c, const float* basis, int rows, float* out, hipStream_t st);
hipError_t launch_project_td(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                             const float* areas, int S, const float* basis, int rows, float* out, unsigned* colmax,
                             hipStream_t st);
// Split-operand form of the projected blocks (BF16 hi / lo, column-major, rows of aperture a
// at a * project_rows_pad(rows) + j) and the SYRK that consumes it.
int project_rows_pad(int rows);
// format 0: BF16 hi / lo; 1: FP16 hi / lo of column s scaled by 2^col_exp[s].
hipError_t launch_project_split(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                                const float* areas, int S, const float* basis, int rows, int format,
                                const float* col_exp, void* hi, void* lo, long long ld, hipStream_t st);
struct SplitShape {
    int tiles, splits, per_split;
};
SplitShape syrk_split_shape(int m, int n, int col0);
size_t syrk_split_workspace(int m, int n, int col0);
hipError_t launch_syrk_split(int format, const void* hi, const void* lo, long long ld, int m, int n, int col0,
                             const float* col_exp, float* out, void* workspace, hipStream_t st);
typedef dsplat_qn_spec QnSpec;
typedef dsplat_joint_layout JointLayout;
hipError_t launch_joint_unpack(int count, const float* vector, const JointLayout& layout, const float* centers,
                               const float* normals, const float* coefficients, const float* root,
                               const float* coefficient_scale, float* dcenters, float* dnormals, float* dweights,
                               hipStream_t st);
hipError_t launch_joint_pack(int count, const float* gcenters, const float* gnormals, const float* gweights,
                             const JointLayout& layout, const float* centers, const float* normals,
                             const float* coefficients, const float* root, const float* coefficient_scale,
                             const float* vector, const float* damping, float normalization, float* out,
                             hipStream_t st);
size_t qn_workspace(int A, int P, int K);
hipError_t launch_qn_search(const Chirp& c, const float* pos, int A, const float2* res, const float* ctr,
                            const float* nrm, const float* areas, int count, const QnSpec& sp, float* oc, float* on,
                            float* ov, void* workspace, hipStream_t st);
// Lockstep ascent on sample-domain chirps: rounds of certificate_gradient evaluations of every
// live candidate and one state-machine step each (qn_search.hip).
size_t qn_lockstep_workspace(const Chirp& c, int A, int count);
hipError_t launch_qn_lockstep(const Chirp& c, const float* pos, int A, const void* operands, const float* ctr,
                              const float* nrm, const float* areas, int count, const QnSpec& sp, float* oc,
                              float* on, float* ov, void* workspace, hipStream_t st);
size_t elimination_workspace(int n, int count);
hipError_t launch_elimination(const double* W, const double* b, int n, int count, const int* limited, int limit,
                              long long* removed, void* workspace, hipStream_t st);
size_t syrk_workspace(int m, int n, int col0, int mode);
hipError_t launch_syrk(const float* rows, int m, int n, int col0, int mode, const unsigned* column_max, float* out,
                       void* workspace, hipStream_t st);

// Splat form (Chirp::form, from the call's dsplat_fmcw): 0 = sample-domain MFMA kernels where
// the chirp allows them (S in {128, 256}), 1 = bin-domain cotangent kernels everywhere.
bool td_supported(const Chirp& c);
bool use_td(const Chirp& c);
struct TdVjpShape {
    int tiles;      // 128-surfel tiles
    int ranges;     // aperture ranges (partials per surfel)
    int per_range;  // apertures per range
};
TdVjpShape td_vjp_shape(int A, int S);
size_t td_vjp_workspace(const Chirp& c, int A, int S);
hipError_t launch_td_render(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                            const float2* w, int S, float2* out, hipStream_t st);
hipError_t launch_td_jvp(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                         const float2* w, const float* dctr, const float* dnrm, const float2* dw, int S, float2* out,
                         hipStream_t st);
// U = DFT of K-bin rows to the S samples (sign +1), the sample-domain form of a residual.
hipError_t launch_dft_rows_up(const Chirp& c, const float2* in, int rows, float2* out, hipStream_t st);
// Sample-domain certificate sums and their gradients (cert_td.hip).
constexpr int kCgValues = 21;  // corr (2), norm, dcorr re / im (6 + 6), dnorm (6)
size_t cert_operands_bytes(const Chirp& c, int A);
hipError_t launch_cert_operands(const Chirp& c, const float2* residual, int A, void* operands, hipStream_t st);
size_t cert_grad_workspace(const Chirp& c, int A, int C);
// live (optional, [C]): candidates with live[s] == 0 may be skipped, their rows of out left as they were.
hipError_t launch_cert_grad(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm, int C,
                            const void* operands, float* out, void* ws, hipStream_t st, const int* live = nullptr);
hipError_t launch_td_vjp(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                         const float2* w, int S, const float2* up, void* ws, float* gc, float* gn, float* gw,
                         hipStream_t st);
hipError_t launch_vjp_reduce(const float* partial, int S, int nblk, float* gc, float* gn, float* gw,
                             hipStream_t st);
// Sample-domain VJP by bin moments (vjp_moments.hip): chirps with 128 or 256 samples and pad >= 2.
bool vjp_moments_supported(const Chirp& c);
size_t vjp_moments_workspace(int A, int S);
hipError_t launch_vjp_moments(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                              const float2* w, int S, const float2* up, void* ws, float* gc, float* gn, float* gw,
                              hipStream_t st);

constexpr int kCholMaxSolve = 1 << 20;  // rows of the blocked solve
constexpr int kPowerMaxN = 16384;      // LDS holds the power-iteration vector
hipError_t launch_cholesky(double* A, int n, int lda, int* info, hipStream_t st);
hipError_t launch_cholesky_f32(float* A, int n, int lda, int* info, hipStream_t st);
hipError_t launch_cholesky_solve(const double* L, int n, int lda, double* B, int nrhs, int ldb, double* W,
                                 hipStream_t st);
size_t cholesky_solve_workspace(int n, int nrhs);
hipError_t launch_power_iteration(const double* h, int n, int ld, int steps, const double* start, double* out,
                                  double* work, hipStream_t st);

hipError_t launch_render(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                         const float2* w, int S, float2* out, hipStream_t st);
hipError_t launch_jvp(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm, const float2* w,
                      const float* dctr, const float* dnrm, const float2* dw, int S, float2* out, hipStream_t st);
hipError_t launch_vjp(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm, const float2* w,
                      int S, const float2* up, float* partial, float* gc, float* gn, float* gw, hipStream_t st);

}  // namespace dsplat
