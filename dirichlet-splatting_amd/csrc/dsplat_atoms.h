// Dictionary atoms generated on chip, column by column (reference: the per-entry atoms of
// surfel.slang:313-378), shared by the fused Gram and the projected-design kernels.
#pragma once

#include "dsplat_device.h"

namespace dsplat {

// Per-column pair state of one aperture (register resident across its K-steps).
template <int P>
struct Column {
    float2 g[P];        // cw nu_r / S per bin class r (cw the column's complex weight)
    float2 lobe_value;  // cw D(-delta)
    float t;            // tan(pi delta / N)
    int lobe;           // nearest bin, -1 if the column is padding
};

// Column state of a pair with complex weight cw (the atom is cw D(k - k0)).
template <int P>
__device__ __forceinline__ Column<P> column_of_pair(const Chirp& c, const Pair& q, float2 cw) {
    Column<P> col;
    col.t = tan_small(c.pi_N * q.delta);
    col.lobe = q.lobe;
    float sb[P], cb[P];
    class_angles<P>(q.delta, q.lobe, sb, cb);
#pragma unroll
    for (int r = 0; r < P; ++r) col.g[r] = c.inv_S * cmul(cw, make_float2(sb[r] * cb[r], -sb[r] * sb[r]));
    float2 D, Dp;
    lobe_direct(c, q.delta, &D, &Dp);
    col.lobe_value = cmul(cw, D);
    return col;
}

template <int P>
__device__ __forceinline__ Column<P> column_state(const Chirp& c, const float* __restrict__ pos, int a,
                                                  const float* __restrict__ ctr, const float* __restrict__ nrm,
                                                  const float* __restrict__ areas, float scale, int s, int S) {
    Column<P> col;
    col.lobe = -1;
    col.t = 0.5f;
    col.lobe_value = make_float2(0.f, 0.f);
#pragma unroll
    for (int r = 0; r < P; ++r) col.g[r] = make_float2(0.f, 0.f);
    if (s >= S) return col;
    const Pair q = pair_geometry(c, pos[3 * a], pos[3 * a + 1], pos[3 * a + 2], ctr[3 * s], ctr[3 * s + 1],
                                 ctr[3 * s + 2], nrm[3 * s], nrm[3 * s + 1], nrm[3 * s + 2]);
    return column_of_pair<P>(c, q, (q.amp * sqrtf(areas[s]) * scale) * q.rot);
}

// Atom of a column at bin k (0 beyond K), cot form with the derivative table (T_0 = 0);
// cls = k mod P is the bin class (class_angles already rotated the classes by the lobe).
template <int P>
__device__ __forceinline__ float2 column_atom(const Chirp& c, const float* __restrict__ T, const Column<P>& col,
                                              int k, int cls) {
    if (k >= c.K) return make_float2(0.f, 0.f);
    const float cot = cot_bin(T[c.N - col.lobe + k], col.t);
    float2 gr = col.g[0];
#pragma unroll
    for (int r = 1; r < P; ++r)
        if (cls == r) gr = col.g[r];
    float2 v = make_float2(gr.x * cot - gr.y, gr.y * cot + gr.x);
    return k == col.lobe ? col.lobe_value : v;
}

}  // namespace dsplat
