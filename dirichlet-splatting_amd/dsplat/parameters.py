"""Per-atom parameter constraints: which components move, the unit of a solver step, and the
feasible set a proposal is pulled back into.

Public surface of the reference's dsplat/parameters.py:15-83 (ParameterConfig with mask,
project, tangent, differential; periodic_delta). A parameter holds one row per atom: a
vector field [S, D] (centres, normals) or a scalar field [S] (areas), which is handled as
[S, 1]. Solver steps are dimensionless; `scale` converts one unit into physical length.
"""

from dataclasses import dataclass

import torch
import torch.nn.functional as F


def _rows(x):
    """A field as [S, D]: scalar fields gain a trailing axis of one."""
    return x.unsqueeze(-1) if x.ndim == 1 else x


def _like(rows, original):
    """Undo _rows for the shape of original."""
    return rows[:, 0] if original.ndim == 1 else rows


def _onto_unit_sphere(value, reference, free):
    """Unit vectors from value [S, D] whose frozen components (free False) are kept: the free
    components are rescaled to the length 1 - |frozen|^2 leaves them. A free part of zero
    length takes its direction from reference; where that is zero too, no direction exists."""
    zero = torch.zeros((), dtype=value.dtype, device=value.device)
    frozen_part = torch.where(free, zero, value)
    free_part = value - frozen_part
    direction = torch.where(free_part.norm(dim=-1, keepdim=True) > 0, free_part, torch.where(free, reference, zero))
    length = direction.norm(dim=-1, keepdim=True)
    budget = (1 - frozen_part.square().sum(-1, keepdim=True)).clamp(min=0).sqrt()
    if bool(((budget > 0) & (length == 0)).any()):
        raise ValueError("Cannot normalize a zero vector without a reference direction.")
    return frozen_part + direction * budget / length.clamp(min=1e-30)


_MASKS: dict = {}


@dataclass(frozen=True)
class ParameterConfig:
    trainable: bool | tuple[bool, ...] = True  # one flag for the field, or one per component
    scale: float = 1.0  # physical length of a unit solver step
    lower: float | tuple[float, ...] | None = None  # box bounds, one or per component
    upper: float | tuple[float, ...] | None = None
    period: tuple[float, ...] | None = None  # wrap-around per component
    normalize: bool = False  # the field holds unit vectors (surfel normals)

    def mask(self, value):
        """bool [D]: the components a step may change (D = 1 for a scalar field). Masks are
        made once per (flags, width, device) and shared read-only, so a solver iteration
        copies nothing from the host (and can be captured in a HIP graph)."""
        width = value.shape[-1] if value.ndim == 2 else 1
        key = (self.trainable, width, str(value.device))
        if key not in _MASKS:
            _MASKS[key] = torch.as_tensor(self.trainable, dtype=torch.bool, device=value.device).expand(width)
        return _MASKS[key]

    def project(self, proposed, reference):
        """The feasible value nearest proposed: wrapped, clamped to the box, frozen components
        taken from reference, unit length restored for a normalized field."""
        free = self.mask(reference)
        value = _rows(proposed)
        if self.period is not None:
            value = torch.remainder(value, value.new_tensor(self.period))
        bounds = (self.lower, torch.maximum), (self.upper, torch.minimum)
        for bound, pick in bounds:
            if bound is not None:
                value = pick(value, value.new_tensor(bound))
        value = torch.where(free, value, _rows(reference))
        if self.normalize and bool(free.any()):
            value = _onto_unit_sphere(value, _rows(reference), free)
        return _like(value, proposed)

    def tangent(self, vector, value):
        """The part of an update [..., S, D] that project keeps at the unit vectors value [S, D]:
        no frozen component, and no radial part within the free components (whose squared
        length is 1 minus that of the frozen ones). Fields that are not normalized pass."""
        if not self.normalize:
            return vector
        free = self.mask(value)
        zero = torch.zeros((), dtype=value.dtype, device=value.device)
        moving = torch.where(free, value, zero)
        budget = (1 - torch.where(free, zero, value).square().sum(-1, keepdim=True)).clamp(min=0)
        step = torch.where(free, vector, zero)
        along = (step * moving).sum(-1, keepdim=True) * moving / budget.clamp(min=1e-30)
        return torch.where(budget > 0, step - along, zero)

    def differential(self, jacobian, value):
        """A dictionary derivative [..., S, D] chained through the retraction (unit fields),
        restricted to the free components and scaled to solver units."""
        if self.normalize:
            jacobian = self.tangent(jacobian, F.normalize(value, dim=-1))
        return self.scale * jacobian[..., self.mask(value)]


def periodic_delta(delta, period):
    """Differences folded per axis into [-period / 2, period / 2]; period None leaves them."""
    if period is None:
        return delta
    span = torch.as_tensor(period, dtype=delta.dtype, device=delta.device)
    return delta - span * torch.round(delta / span)
