// The parameter maps of the joint Gauss-Newton operator around the splat's JVP and VJP
// (reference: refinement.py:136-146 unpack, 148-158 adjoint, 170-174 matvec), one thread
// per atom, so that an operator application is the JVP, the VJP and two small launches
// instead of some fifty tensor operations.
//
// The unknowns are dimensionless: per trainable field a section of count x D values (D = 3
// for centres and normals, 1 for areas) in solver units, then the coefficients (one real
// value per atom, or (Re, Im) pairs). Every component of a field is free on this path.
//   unpack:  field tangent = section * scale, for a unit field without its radial part
//            (parameters.py:72-83 tangent with every component free: t - (t . n) n); the
//            weight tangent of w = b sqrt(area): db sqrt(area) + b d(area) / (2 sqrt(area)),
//            db = the coefficient unknowns times coefficient_scale.
//   pack:    the VJP's gradients through the same maps back to the unknowns (the area's as
//            Re(conj(g_w) b) / (2 sqrt(area)), the coefficients' as g_w sqrt(area) times
//            coefficient_scale), then out = pack / normalization + damping * vector.
// Floating-point contraction is off: each value is rounded where the tensor form rounds it.
#include "dsplat_internal.h"

#pragma clang fp contract(off)

namespace dsplat {

// t - (t . n) n, the dot product summed left to right
__device__ __forceinline__ void unit_tangent(float t[3], const float* n) {
    const float n0 = n[0], n1 = n[1], n2 = n[2];
    const float dot = (t[0] * n0 + t[1] * n1) + t[2] * n2;
    t[0] = t[0] - dot * n0;
    t[1] = t[1] - dot * n1;
    t[2] = t[2] - dot * n2;
}

__global__ __launch_bounds__(256) void joint_unpack_kernel(int count, const float* __restrict__ v, JointLayout L,
                                                           const float* __restrict__ centers,
                                                           const float* __restrict__ normals,
                                                           const float2* __restrict__ b, const float* __restrict__ root,
                                                           const float* __restrict__ cscale, float* __restrict__ dc,
                                                           float* __restrict__ dn, float2* __restrict__ dw) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= count) return;
    auto field = [&](int offset, float scale, int unit, const float* value, float* out) {
        float t[3] = {0.0f, 0.0f, 0.0f};
        if (offset >= 0) {
#pragma unroll
            for (int j = 0; j < 3; ++j) t[j] = v[offset + 3 * s + j] * scale;
            if (unit) unit_tangent(t, value + 3 * s);
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) out[3 * s + j] = t[j];
    };
    field(L.centers, L.center_scale, L.centers_unit, centers, dc);
    field(L.normals, L.normal_scale, L.normals_unit, normals, dn);
    const float da = L.areas >= 0 ? v[L.areas + s] * L.area_scale : 0.0f;
    const float c = cscale[s];
    const float2 db = L.complex_mode ? make_float2(v[L.tail + 2 * s] * c, v[L.tail + 2 * s + 1] * c)
                                     : make_float2(v[L.tail + s] * c, 0.0f);
    const float r = root[s];
    const float2 bb = b[s];
    const float two_r = 2.0f * r;
    dw[s] = make_float2(db.x * r + bb.x * da / two_r, db.y * r + bb.y * da / two_r);
}

__global__ __launch_bounds__(256) void joint_pack_kernel(int count, const float* __restrict__ gc,
                                                         const float* __restrict__ gn, const float2* __restrict__ gw,
                                                         JointLayout L, const float* __restrict__ centers,
                                                         const float* __restrict__ normals,
                                                         const float2* __restrict__ b, const float* __restrict__ root,
                                                         const float* __restrict__ cscale, const float* __restrict__ v,
                                                         const float* __restrict__ damping, float normalization,
                                                         float* __restrict__ out) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= count) return;
    auto put = [&](int i, float value) { out[i] = value / normalization + damping[i] * v[i]; };
    auto field = [&](int offset, float scale, int unit, const float* value, const float* grad) {
        if (offset < 0) return;
        float t[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) t[j] = grad[3 * s + j];
        if (unit) unit_tangent(t, value + 3 * s);
#pragma unroll
        for (int j = 0; j < 3; ++j) put(offset + 3 * s + j, t[j] * scale);
    };
    field(L.centers, L.center_scale, L.centers_unit, centers, gc);
    field(L.normals, L.normal_scale, L.normals_unit, normals, gn);
    const float r = root[s];
    const float2 g = gw[s];
    if (L.areas >= 0) {
        const float2 bb = b[s];
        put(L.areas + s, ((g.x * bb.x + g.y * bb.y) / (2.0f * r)) * L.area_scale);
    }
    const float c = cscale[s];
    if (L.complex_mode) {
        put(L.tail + 2 * s, (g.x * r) * c);
        put(L.tail + 2 * s + 1, (g.y * r) * c);
    } else {
        put(L.tail + s, (g.x * r) * c);
    }
}

hipError_t launch_joint_unpack(int count, const float* vector, const JointLayout& layout, const float* centers,
                               const float* normals, const float* coefficients, const float* root,
                               const float* coefficient_scale, float* dcenters, float* dnormals, float* dweights,
                               hipStream_t st) {
    hipLaunchKernelGGL(joint_unpack_kernel, dim3((count + 255) / 256), dim3(256), 0, st, count, vector, layout,
                       centers, normals, reinterpret_cast<const float2*>(coefficients), root, coefficient_scale,
                       dcenters, dnormals, reinterpret_cast<float2*>(dweights));
    return hipGetLastError();
}

hipError_t launch_joint_pack(int count, const float* gcenters, const float* gnormals, const float* gweights,
                             const JointLayout& layout, const float* centers, const float* normals,
                             const float* coefficients, const float* root, const float* coefficient_scale,
                             const float* vector, const float* damping, float normalization, float* out,
                             hipStream_t st) {
    hipLaunchKernelGGL(joint_pack_kernel, dim3((count + 255) / 256), dim3(256), 0, st, count, gcenters, gnormals,
                       reinterpret_cast<const float2*>(gweights), layout, centers, normals,
                       reinterpret_cast<const float2*>(coefficients), root, coefficient_scale, vector, damping,
                       normalization, out);
    return hipGetLastError();
}

}  // namespace dsplat
