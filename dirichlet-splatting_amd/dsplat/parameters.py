"""Constraints of per-atom parameters: free components, step scales and feasible sets.

Same interface as dsplat/parameters.py of the reference (parameters.py:15-83). A
parameter is a tensor with one row per atom (centers [S, 3], areas [S]); solver steps
are dimensionless and go through ParameterConfig: `scale` is the physical length of a
unit step, `project` restores feasibility, frozen components keep their values.
"""

from dataclasses import dataclass

import torch
import torch.nn.functional as F


@dataclass(frozen=True)
class ParameterConfig:
    trainable: bool | tuple[bool, ...] = True  # one flag, or one per component
    scale: float = 1.0
    lower: float | tuple[float, ...] | None = None
    upper: float | tuple[float, ...] | None = None
    period: tuple[float, ...] | None = None
    normalize: bool = False  # unit vectors, such as surfel normals

    def mask(self, value):
        """Boolean trainable mask [D] (D = 1 for a vector parameter)."""
        flags = torch.as_tensor(self.trainable, dtype=torch.bool, device=value.device)
        return flags.expand(value.shape[1] if value.ndim == 2 else 1)

    def project(self, proposed, reference):
        """Feasible values nearest to `proposed`; frozen components come from `reference`.

        With normalize, the free components are rescaled so the vector has unit length,
        taking the room the frozen components leave; a zero free part takes the
        reference's direction (parameters.py:29-55).
        """
        vector = proposed.ndim == 1
        value = proposed[:, None] if vector else proposed
        if self.period is not None:
            value = value.remainder(value.new_tensor(self.period))
        if self.lower is not None:
            value = torch.maximum(value, value.new_tensor(self.lower))
        if self.upper is not None:
            value = torch.minimum(value, value.new_tensor(self.upper))
        ref = reference[:, None] if vector else reference
        mask = self.mask(reference)[None]
        value = torch.where(mask, value, ref)
        if self.normalize and bool(mask.any()):
            fixed = torch.where(mask, 0, value)
            free = value - fixed
            has_length = free.norm(dim=-1, keepdim=True) > 0
            free = torch.where(has_length, free, torch.where(mask, ref, 0))
            room = (1 - fixed.square().sum(-1, keepdim=True)).clamp(min=0).sqrt()
            length = free.norm(dim=-1, keepdim=True)
            if bool(((room > 0) & (length == 0)).any()):
                raise ValueError("Cannot normalize a zero vector without a reference direction.")
            value = fixed + free * room / length.clamp(min=1e-30)
        return value[:, 0] if vector else value

    def tangent(self, vector, value):
        """The part of an update [..., S, D] that project keeps at a unit vector `value`:
        frozen components drop out, and so does the radial part within the free ones
        (parameters.py:57-70)."""
        if not self.normalize:
            return vector
        mask = self.mask(value)
        free = torch.where(mask, value, 0)
        room = (1 - torch.where(mask, 0, value).square().sum(-1, keepdim=True)).clamp(min=0)
        vector = torch.where(mask, vector, 0)
        radial = (vector * free).sum(-1, keepdim=True) * free / room.clamp(min=1e-30)
        return torch.where(room > 0, vector - radial, 0)

    def differential(self, jacobian, value):
        """Chain a dictionary derivative [..., S, D] through the retraction, mask and scale."""
        if self.normalize:
            jacobian = self.tangent(jacobian, F.normalize(value, dim=-1))
        return jacobian[..., self.mask(value)] * self.scale


def periodic_delta(delta, period):
    """Differences folded into [-period/2, period/2] per axis; None leaves them unchanged."""
    if period is None:
        return delta
    period = torch.as_tensor(period, dtype=delta.dtype, device=delta.device)
    return delta - period * torch.round(delta / period)
