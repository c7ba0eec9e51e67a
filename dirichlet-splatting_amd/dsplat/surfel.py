"""Oriented surfels: the scattering primitive of the physical measurements.

Same functions as dsplat/surfel.py:45-129 of the reference: the sqrt(area) chain
around the surfel kernels. A surfel has centers [S, 3] (m), outward unit normals
[S, 3] and areas [S] (m^2); its reflectivity b is the linear coefficient. Seen from
a monostatic sensor at p, with r = |p - c|, u = (p - c) / r and cos = u . n:

    amplitude = sqrt(area) cos^2 F(cos) lambda / (8 pi r)   zero when cos <= 0
    F(cos)    = (root - cos) / (root + cos),  root = sqrt(fresnel_n^2 - 1 + cos^2)

and the chirp adds the phase 2 pi (fc tau + K tau (t0 - tau / 2)), tau = 2 r / c, at
range bin r / spacing.
"""

from .kernels import surfel as kernel


def render(config, positions, params, coefficients, fresnel_n=0.0):
    """Range profiles [A, K] of surfels with reflectivities coefficients; differentiable
    (surfel.py:45-50)."""
    weights = coefficients * params["areas"].sqrt()
    return kernel.render(config, positions, params["centers"], params["normals"], weights, fresnel_n)


def jvp(config, positions, params, coefficients, tangent, delta_coefficients, fresnel_n=0.0):
    """Change of the profiles along tangent (centers, normals, areas) and delta_coefficients
    (surfel.py:53-61)."""
    scale = params["areas"].sqrt()
    weights = coefficients * scale
    delta_weights = delta_coefficients * scale + coefficients * tangent["areas"] / (2 * scale)
    return kernel.jvp(
        config, positions, params["centers"], params["normals"], weights,
        tangent["centers"], tangent["normals"], delta_weights, fresnel_n,
    )  # fmt: skip


def vjp(config, positions, params, coefficients, upstream, fresnel_n=0.0):
    """Gradients of Re(conj(upstream) . profiles): per parameter, and for the coefficients
    (surfel.py:64-77)."""
    scale = params["areas"].sqrt()
    centers, normals, weights = kernel.vjp(
        config, positions, params["centers"], params["normals"], coefficients * scale, upstream, fresnel_n
    )
    areas = (weights.conj() * coefficients).real / (2 * scale)
    return dict(centers=centers, normals=normals, areas=areas), weights * scale
