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


def design(config, positions, params, fresnel_n=0.0, real=False):
    """Dictionary columns of every surfel for whole apertures (surfel.py:80-84)."""
    return kernel.design(config, positions, params["centers"], params["normals"], params["areas"], fresnel_n, real)


def jacobian(config, positions, params, fresnel_n=0.0):
    """Derivatives of the columns [A, K, S] by centers and normals [..., 3] and areas [..., 1]
    (surfel.py:87-97)."""
    columns, dc, dn = kernel.jacobian(config, positions, params["centers"], params["normals"], fresnel_n)
    scale = params["areas"].sqrt()
    return {
        "centers": dc * scale[:, None],
        "normals": dn * scale[:, None],
        "areas": (columns / (2 * scale))[..., None],
    }


def correlate(config, positions, params, residual, fresnel_n=0.0):
    """a^H r and |a|^2 of every surfel's column against the residual profiles [A, K]
    (surfel.py:100-105)."""
    corr, norm = kernel.correlate(config, positions, params["centers"], params["normals"], residual, fresnel_n)
    return corr * params["areas"].sqrt(), norm * params["areas"]


def correlate_gradient(config, positions, params, residual, fresnel_n=0.0, include_normals=False):
    """correlate and its derivatives by centers (and normals): corr, norm, dcorr, dnorm
    (surfel.py:108-120)."""
    corr, norm, dcorr, dnorm = kernel.correlate_gradient(
        config, positions, params["centers"], params["normals"], residual, fresnel_n, include_normals
    )
    scale, area = params["areas"].sqrt(), params["areas"]
    return corr * scale, norm * area, dcorr * scale[:, None], dnorm * area[:, None]


def correlate_orientations(config, positions, params, bank, residual, fresnel_n=0.0, table=None):
    """correlate for each surfel's own normal and every bank normal, [S, 1 + bank]
    (surfel.py:123-129)."""
    corr, norm = kernel.correlate_orientations(
        config, positions, params["centers"], params["normals"], bank, residual, fresnel_n, table
    )
    scale = params["areas"].sqrt()[:, None]
    return corr * scale, norm * scale.square()


def orient_select(config, positions, params, bank, residual, fresnel_n=0.0, table=None, real_part=False):
    """Best of each surfel's own normal and the bank by |s|^2 / |a|^2 (area included), in
    one launch: (value [S] float64, index [S] int64, 0 = own normal, 1 + b = bank b)."""
    return kernel.orient_select(
        config, positions, params["centers"], params["normals"], params["areas"], bank, residual, fresnel_n, table,
        real_part,
    )  # fmt: skip
