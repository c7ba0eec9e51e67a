"""Oriented surfels, the scattering primitive: the area chain between the surfel parameters
and the HIP splat kernels.

Functions of the reference's dsplat/surfel.py:45-129. A surfel is centres [S, 3] (metres),
outward unit normals [S, 3] and areas [S] (square metres); its reflectivity b is the linear
coefficient. From a monostatic sensor at p, with r = |p - c|, u = (p - c) / r, cos = u . n:

    amplitude = sqrt(area) cos^2 F(cos) lambda / (8 pi r),  zero when cos <= 0
    F(cos)    = (root - cos) / (root + cos),  root = sqrt(fresnel_n^2 - 1 + cos^2)

and the chirp adds the phase 2 pi (fc tau + K tau (t0 - tau / 2)), tau = 2 r / c, at range
bin r / spacing. The kernels take one complex weight per surfel, w = b sqrt(area); this
module maps (b, area) to w and chains every derivative back through it.
"""

import torch

from .kernels import surfel as kernel


class _Footprint:
    """sqrt(area) of every surfel, and the chain rule of w = b sqrt(area)."""

    def __init__(self, params):
        self.area = params["areas"]
        self.root = self.area.sqrt()

    def weights(self, coefficients):
        return coefficients * self.root

    def weight_tangent(self, coefficients, delta_coefficients, delta_area):
        # dw = db sqrt(area) + b d(area) / (2 sqrt(area))
        return delta_coefficients * self.root + coefficients * delta_area / (2 * self.root)

    def area_gradient(self, weight_gradient, coefficients):
        # dL/d(area) = Re(conj(dL/dw) b) / (2 sqrt(area))
        return (weight_gradient.conj() * coefficients).real / (2 * self.root)

    def per_column(self, values):
        """values [..., S, w] times sqrt(area) per surfel column."""
        return values * self.root[:, None]


def _geometry(params):
    return params["centers"], params["normals"]


def render(config, positions, params, coefficients, fresnel_n=0.0):
    """Range profiles [A, K] of surfels with reflectivities b = coefficients; differentiable
    (surfel.py:45-50)."""
    return kernel.render(config, positions, *_geometry(params), _Footprint(params).weights(coefficients), fresnel_n)


def jvp(config, positions, params, coefficients, tangent, delta_coefficients, fresnel_n=0.0):
    """Directional derivative of the profiles along tangent (centers, normals, areas) and
    delta_coefficients (surfel.py:53-61)."""
    foot = _Footprint(params)
    return kernel.jvp(
        config, positions, *_geometry(params), foot.weights(coefficients), tangent["centers"], tangent["normals"],
        foot.weight_tangent(coefficients, delta_coefficients, tangent["areas"]), fresnel_n,
    )  # fmt: skip


def vjp(config, positions, params, coefficients, upstream, fresnel_n=0.0):
    """Gradients of Re(conj(upstream) . profiles): a dict per parameter, and the coefficients'
    (surfel.py:64-77)."""
    foot = _Footprint(params)
    g_centers, g_normals, g_weights = kernel.vjp(
        config, positions, *_geometry(params), foot.weights(coefficients), upstream, fresnel_n
    )
    grads = dict(centers=g_centers, normals=g_normals, areas=foot.area_gradient(g_weights, coefficients))
    return grads, foot.weights(g_weights)


def design(config, positions, params, fresnel_n=0.0, real=False):
    """Dictionary columns (sqrt(area) included) of every surfel for whole apertures
    (surfel.py:80-84); see kernels.surfel.design for the layouts."""
    return kernel.design(config, positions, *_geometry(params), params["areas"], fresnel_n, real)


def jacobian(config, positions, params, fresnel_n=0.0):
    """Derivatives of the columns [A, K, S] by centres and normals [..., 3] and by the area
    [..., 1] (surfel.py:87-97)."""
    columns, by_center, by_normal = kernel.jacobian(config, positions, *_geometry(params), fresnel_n)
    foot = _Footprint(params)
    return {
        "centers": foot.per_column(by_center),
        "normals": foot.per_column(by_normal),
        "areas": (columns / (2 * foot.root))[..., None],
    }


def correlate(config, positions, params, residual, fresnel_n=0.0):
    """a^H r and |a|^2 of every surfel's column a against the residual profiles [A, K]
    (surfel.py:100-105)."""
    corr, norm = kernel.correlate(config, positions, *_geometry(params), residual, fresnel_n)
    foot = _Footprint(params)
    return corr * foot.root, norm * foot.area


def adjoint_correlation(config, positions, params, residual, fresnel_n=0.0):
    """A^H r of every surfel's column (sqrt(area) included): correlate's first output, as the
    coefficient gradient of Re<r, A b> at b = 1, which the VJP kernel forms on the matrix cores
    in the sample domain (kernels/surfel.py:123-137 run for kernels/surfel.py:242-260)."""
    ones = params["areas"].new_ones(len(params["areas"]), dtype=torch.complex64)
    return vjp(config, positions, params, ones, residual, fresnel_n)[1]


def correlate_gradient(config, positions, params, residual, fresnel_n=0.0, include_normals=False):
    """correlate and its derivatives by centres (and normals): corr, norm, dcorr, dnorm
    (surfel.py:108-120)."""
    corr, norm, dcorr, dnorm = kernel.correlate_gradient(
        config, positions, *_geometry(params), residual, fresnel_n, include_normals
    )
    foot = _Footprint(params)
    return corr * foot.root, norm * foot.area, foot.per_column(dcorr), dnorm * foot.area[:, None]


def correlate_orientations(config, positions, params, bank, residual, fresnel_n=0.0, table=None):
    """correlate for each surfel's own normal and every bank normal, [S, 1 + B]
    (surfel.py:123-129)."""
    corr, norm = kernel.correlate_orientations(config, positions, *_geometry(params), bank, residual, fresnel_n, table)
    root = _Footprint(params).root[:, None]
    return corr * root, norm * root.square()


def orient_select(config, positions, params, bank, residual, fresnel_n=0.0, table=None, real_part=False):
    """Best of each surfel's own normal and the bank by |s|^2 / |a|^2 (area included), in one
    launch: (value [S] float64, index [S] int64; 0 = own normal, 1 + b = bank normal b)."""
    return kernel.orient_select(
        config, positions, *_geometry(params), params["areas"], bank, residual, fresnel_n, table, real_part
    )
