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
CorShape cor_shape(int A, int S);
size_t cor_smem(const Chirp& c);

hipError_t launch_design(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                         const float* areas, int S, int complex_layout, float* out, hipStream_t st);
hipError_t launch_jacobian(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm, int S,
                           float2* col, float2* jac, hipStream_t st);
hipError_t launch_correlate(const Chirp& c, int W, const float* pos, int A, const float* ctr, const float* nrm, int S,
                            const float2* res, float* partial, float* out, hipStream_t st);
hipError_t launch_orientations(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                               const float* bank, int B, int S, const float2* res, const float2* table,
                               const float* norm_table, int oversample, float* partial, float* out, hipStream_t st);

hipError_t launch_render(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm,
                         const float2* w, int S, float2* out, hipStream_t st);
hipError_t launch_jvp(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm, const float2* w,
                      const float* dctr, const float* dnrm, const float2* dw, int S, float2* out, hipStream_t st);
hipError_t launch_vjp(const Chirp& c, const float* pos, int A, const float* ctr, const float* nrm, const float2* w,
                      int S, const float2* up, float* partial, float* gc, float* gn, float* gw, hipStream_t st);

}  // namespace dsplat
