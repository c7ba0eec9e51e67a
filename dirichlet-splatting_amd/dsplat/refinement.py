"""Local geometry refinement: dense reduced-VarPro Levenberg-Marquardt and matrix-free joint
Gauss-Newton.

Public surface of the reference's dsplat/refinement.py:13-227 (trainable, reduced_system,
dense_slide, joint_system, joint_slide). Both refinements propose a move of every trainable
atom and keep it only if the refitted objective is lower; steps are dimensionless, in units
of each ParameterConfig.scale.

  dense  the Gauss-Newton normal matrix of the reduced objective (coefficients eliminated
         exactly) on sampled observation rows, solved densely with diagonal damping
  joint  geometry and coefficients together: J^H J as a JVP followed by a VJP, solved by
         Jacobi-preconditioned CG on an aperture subset; the diagonal comes from Rademacher
         probes
"""

import torch

from .varpro import accept, fit_coefficients, pcg, screen_bar, solve


def trainable(params, specs):
    """(key, spec) of every field with at least one free component."""
    return [(key, spec) for key, spec in specs.items() if spec.mask(params[key]).any()]


def _moved(params, layout, delta, fraction):
    """params with each trainable field moved by fraction * scale * its section of delta
    [S, free components] and projected back onto its constraints."""
    proposal = {key: value.clone() for key, value in params.items()}
    column = 0
    for key, spec in layout:
        value = proposal[key]
        grid = value[:, None] if value.ndim == 1 else value.clone()
        free = spec.mask(value)
        width = int(free.sum())
        grid[:, free] += fraction * spec.scale * delta[:, column : column + width]
        proposal[key] = spec.project(grid[:, 0] if value.ndim == 1 else grid, params[key])
        column += width
    return proposal


@torch.no_grad()
def reduced_system(model, params, target, specs, config, rows):
    """Gauss-Newton normal matrix and gradient of the ridge-VarPro objective on the given rows,
    the coefficients' exact differential eliminated (refinement.py:20-52)."""
    a = model.design(params, rows)
    y = target.flatten()[rows]
    m = len(a)
    gram, rhs = a.H @ a / m, a.H @ y / m
    b = fit_coefficients(gram, rhs, config)
    residual = a @ b.to(a.dtype) - y
    layout = trainable(params, specs)
    atoms = torch.arange(len(b), device=a.device)
    raw = model.jacobian(params, rows, atoms, [key for key, _ in layout])
    d_atoms = torch.cat([spec.differential(raw[key], params[key]) for key, spec in layout], -1)
    h = (gram.real if config.real else gram) + config.ridge * torch.eye(len(b), dtype=b.dtype, device=a.device)
    j = (d_atoms * b[None, :, None]).reshape(m, -1)
    cross = a.H @ j / m
    columns = torch.arange(j.shape[1], device=a.device)
    cross[atoms.repeat_interleave(d_atoms.shape[-1]), columns] += (d_atoms.conj() * residual[:, None, None]).sum(0).flatten() / m
    d_b = -solve(h, cross.real if config.real else cross)
    reduced = j + a @ d_b.to(a.dtype)
    normal = (reduced.H @ reduced / m + config.ridge * d_b.H @ d_b).real
    gradient = (reduced.H @ residual / m + config.ridge * d_b.H @ b).real
    return normal, gradient


def dense_slide(state, config, screen=None):
    """Damped LM steps on lm_rows sampled rows with a per-atom trust radius and halving
    backtracks (refinement.py:55-98); returns the state and the accepted steps."""
    layout = trainable(state.params, state.parameters)
    rows = state.model.sample_rows(config.lm_rows, config.seed + state.iteration, state.target.device)
    accepted = 0
    for _ in range(config.lm_steps):
        normal, gradient = reduced_system(state.model, state.params, state.target, state.parameters, config, rows)
        scale = normal.diag().clamp(min=normal.diag().max() * 1e-6 + 1e-30)
        delta = solve(normal + config.damping * torch.diag(scale), -gradient).reshape(len(state.coefficients), -1)
        delta *= (config.trust_radius / delta.norm(dim=1, keepdim=True).clamp(min=1e-30)).clamp(max=1)
        if not torch.isfinite(delta).all():
            break
        bar = screen_bar(screen, state.params, config)
        for fraction in (1.0, 0.5, 0.25):
            trial = accept(state, _moved(state.params, layout, delta, fraction), config, bar)
            if trial is not None:
                state, accepted = trial, accepted + 1
                break
        else:
            break
    return state, accepted


class _JointSystem:
    """The scaled Gauss-Newton system of every trainable geometry component and every
    coefficient on one model (refinement.py:104-190 of the reference).

    Unknowns are dimensionless: geometry in units of each spec's scale, coefficients relative
    to max(|b|, 0.1 RMS |b|); the data term is normalized by M |y|^2 / M (the target power)."""

    def __init__(self, model, params, coefficients, target, specs, config, seed):
        self.model, self.params, self.specs, self.config = model, params, specs, config
        self.b = coefficients.to(target.dtype)
        self.complex_mode = not config.real
        count = len(self.b)
        self.count = count
        # (key, free components as indices or None for all, slice of the unknowns): the
        # indices are formed once, so no solver iteration waits on a boolean mask
        self.sections, offset = [], 0
        for key, spec in trainable(params, specs):
            free = spec.mask(params[key])
            columns = None if bool(free.all()) else free.nonzero().flatten()
            width = count * int(free.sum())
            self.sections.append((key, columns, slice(offset, offset + width)))
            offset += width
        self.geometry = offset
        self.size = offset + (2 * count if self.complex_mode else count)
        self.amplitude = self.b.abs()
        self.coefficient_scale = self.amplitude.clamp(min=self.amplitude.square().mean().sqrt() * 0.1 + 1e-20)
        power = target.abs().square().mean().clamp(min=1e-30)
        self.normalization = target.numel() * power
        ridge = config.ridge / power
        self.ridge_diagonal = self.amplitude.new_zeros(self.size)
        self.ridge_gradient = self.amplitude.new_zeros(self.size)
        tail = slice(self.geometry, None)
        if self.complex_mode:
            self.ridge_diagonal[tail] = (ridge * self.coefficient_scale.square()).repeat_interleave(2)
            self.ridge_gradient[tail] = torch.view_as_real(self.b * (ridge * self.coefficient_scale)).flatten()
        else:
            self.ridge_diagonal[tail] = ridge * self.coefficient_scale.square()
            self.ridge_gradient[tail] = ridge * self.coefficient_scale * self.b.real
        self.gradient = self.adjoint(model.render(params, self.b) - target) / self.normalization + self.ridge_gradient
        self.diagonal = self._probe_diagonal(target, seed)
        self.damping = config.damping * self.diagonal
        self._fused = self._fused_maps()

    def _fused_maps(self):
        """The unpack / adjoint maps of matvec as two kernels around the splat's JVP and VJP
        (kernels.solver.joint_unpack / joint_pack), when every trainable field is a fully free
        centres, normals or areas field of a splat model on float32 device tensors; None
        otherwise (matvec then runs the tensor form)."""
        from .kernels.solver import joint_layout

        model, params = self.model, self.params
        if not (self.b.is_cuda and self.b.dtype == torch.complex64 and hasattr(model, "sensor")
                and hasattr(model, "fresnel_n") and {"centers", "normals", "areas"} <= params.keys()):  # fmt: skip
            return None
        for key in ("centers", "normals", "areas"):
            value = params[key]
            if value.dtype != torch.float32 or not value.is_contiguous():
                return None
            if value.shape != ((self.count,) if key == "areas" else (self.count, 3)):
                return None
        fields = {}
        for key, columns, section in self.sections:
            spec = self.specs[key]
            if columns is not None or key not in ("centers", "normals", "areas") or (key == "areas" and spec.normalize):
                return None
            fields[key] = (section.start, spec.scale, spec.normalize)
        root = params["areas"].sqrt()
        return {
            "layout": joint_layout(self.geometry, self.complex_mode, **fields),
            "root": root,
            "weights": self.b * root,
            "damping": self.ridge_diagonal + self.damping,
            "normalization": float(self.normalization),
        }

    def _probe_diagonal(self, target, seed):
        """Jacobi diagonal of J^H J from cg_probes Rademacher probes (complex, unit modulus
        parts), plus the ridge, floored at 1e-6 of its largest entry."""
        generator = torch.Generator(device=self.b.device).manual_seed(seed)
        total = torch.zeros_like(self.gradient)
        for _ in range(self.config.cg_probes):
            signs = torch.randint(0, 2, (*target.shape, 2), device=self.b.device, generator=generator)
            probe = torch.view_as_complex((2 * signs - 1).to(target.real.dtype))
            total += self.adjoint(probe).square() / self.normalization
        diagonal = total / self.config.cg_probes + self.ridge_diagonal
        return diagonal.clamp(min=diagonal.max().clamp(min=1e-30) * 1e-6)

    @staticmethod
    def _free(grid, columns):
        return grid if columns is None else grid[:, columns]

    def unpack(self, vector):
        """A dimensionless vector as (parameter tangents, coefficient change)."""
        tangent = {key: torch.zeros_like(value) for key, value in self.params.items()}
        for key, columns, section in self.sections:
            value = tangent[key]
            grid = value[:, None] if value.ndim == 1 else value
            block = vector[section].reshape(self.count, -1) * self.specs[key].scale
            if columns is None:
                grid.copy_(block)
            else:
                grid[:, columns] = block
            tangent[key] = self.specs[key].tangent(value, self.params[key])
        tail = vector[self.geometry :]
        if self.complex_mode:
            return tangent, torch.complex(tail[0::2], tail[1::2]) * self.coefficient_scale
        return tangent, (tail * self.coefficient_scale).to(self.b.dtype)

    def adjoint(self, upstream):
        """J^H upstream in the dimensionless unknowns (the model's VJP, chained)."""
        grads, by_coefficient = self.model.vjp(self.params, self.b, upstream)
        out = self.amplitude.new_zeros(self.size)
        for key, columns, section in self.sections:
            value = self.specs[key].tangent(grads[key], self.params[key])
            grid = value[:, None] if value.ndim == 1 else value
            out[section] = (self._free(grid, columns) * self.specs[key].scale).flatten()
        tail = by_coefficient * self.coefficient_scale
        out[self.geometry :] = torch.view_as_real(tail).flatten() if self.complex_mode else tail.real
        return out

    def matvec(self, vector):
        """(J^H J / normalization + ridge + damping) vector: the model's JVP, then its VJP."""
        if self._fused is not None:
            return self._matvec_fused(vector)
        return (self.adjoint(self.model.jvp(self.params, self.b, *self.unpack(vector))) / self.normalization
                + (self.ridge_diagonal + self.damping) * vector)  # fmt: skip

    def _matvec_fused(self, vector):
        """matvec as unpack kernel, JVP, VJP and pack kernel (four to six launches)."""
        from .kernels import surfel as splat
        from .kernels.solver import joint_pack, joint_unpack

        f = self._fused
        centers, normals = self.params["centers"], self.params["normals"]
        config, positions = self.model.sensor
        dc, dn, dw = joint_unpack(vector, f["layout"], centers, normals, self.b, f["root"], self.coefficient_scale)
        image = splat.jvp(config, positions, centers, normals, f["weights"], dc, dn, dw, self.model.fresnel_n)
        gc, gn, gw = splat.vjp(config, positions, centers, normals, f["weights"], image, self.model.fresnel_n)
        return joint_pack(gc, gn, gw, f["layout"], centers, normals, self.b, f["root"], self.coefficient_scale, vector,
                          f["damping"], f["normalization"])  # fmt: skip

    def proposal(self, vector, fraction):
        """The geometry moved by fraction of vector, the whole move first shortened so that no
        atom leaves the trust radius, then projected onto the constraints."""
        tangent, _ = self.unpack(vector)
        length2 = self.amplitude.new_zeros(self.count)
        for key, columns, _ in self.sections:
            grid = tangent[key][:, None] if tangent[key].ndim == 1 else tangent[key]
            length2 += (self._free(grid, columns) / self.specs[key].scale).square().sum(1)
        shrink = (self.config.trust_radius / length2.sqrt().max().clamp(min=1e-30)).clamp(max=1)
        return {key: self.specs[key].project(value + tangent[key] * (shrink * fraction), value)
                for key, value in self.params.items()}  # fmt: skip


def joint_system(model, params, coefficients, target, specs, config, seed):
    """(matvec, gradient, damped Jacobi diagonal, proposal(vector, fraction)) of the scaled
    joint Gauss-Newton system (refinement.py:104-190)."""
    system = _JointSystem(model, params, coefficients, target, specs, config, seed)
    return system.matvec, system.gradient, system.diagonal + system.damping, system.proposal


@torch.no_grad()
def joint_slide(state, config, screen=None):
    """PCG steps of the joint system on joint_apertures sampled apertures, each accepted on
    every observation with halving backtracks (refinement.py:193-227)."""
    model, target = state.model, state.target
    if 0 < config.joint_apertures < model.apertures:
        ids, model = model.sample_apertures(config.joint_apertures, config.seed + state.iteration)
        target = target[ids]
    accepted = 0
    for inner in range(config.lm_steps):
        matvec, gradient, diagonal, proposal = joint_system(model, state.params, state.coefficients, target,
                                                            state.parameters, config,
                                                            config.seed + state.iteration + inner)  # fmt: skip
        # the batched form runs the iteration's algebra as one kernel and replays the
        # iteration from a HIP graph
        device_loop = getattr(config, "batched", False) and gradient.is_cuda
        delta = pcg(matvec, -gradient, diagonal, iterations=config.cg_iterations, tolerance=config.cg_tolerance,
                    fused=device_loop, graph=device_loop)  # fmt: skip
        if not (torch.isfinite(delta).all() and delta @ gradient < 0):
            break
        bar = screen_bar(screen, state.params, config)
        for fraction in (1.0, 0.5, 0.25, 0.125):
            trial = accept(state, proposal(delta, fraction), config, bar)
            if trial is not None:
                state, accepted = trial, accepted + 1
                break
        else:
            break
    return state, accepted
