"""Host algebra of the solver on the CPU: no kernel runs here.

Known answers and brute force, following the reference's solver tests
(tests/test_solver.py:46-136, 351-406).
"""

import pytest
import torch

from dsplat import ParameterConfig
from dsplat.replacement import backward_elimination, regularized_inverse, utility
from dsplat.varpro import pcg, solve


def test_solve_of_a_singular_system_predicts_like_the_pseudoinverse():
    g = torch.Generator().manual_seed(0)
    basis = torch.randn(6, 3, dtype=torch.float64, generator=g)
    h = basis @ basis.T
    q = basis @ torch.randn(3, dtype=torch.float64, generator=g)
    x = solve(h, q)
    torch.testing.assert_close(basis.T @ x, basis.T @ (torch.linalg.pinv(h, hermitian=True) @ q), atol=1e-8, rtol=0)
    torch.testing.assert_close(h @ x, q, atol=1e-8, rtol=0)
    assert torch.count_nonzero(solve(torch.zeros(3, 3, dtype=torch.float64), torch.zeros(3, dtype=torch.float64))) == 0


def test_backward_elimination_follows_brute_force_refits():
    g = torch.Generator().manual_seed(4)
    x = torch.randn(20, 8, generator=g, dtype=torch.float64)
    h, q = x.T @ x, x.T @ torch.randn(20, generator=g, dtype=torch.float64)

    def objective(keep):
        keep = list(keep)
        return float(-q[keep] @ torch.linalg.solve(h[keep][:, keep], q[keep]))

    alive, expected = set(range(8)), []
    for _ in range(3):
        best = min(alive, key=lambda j: objective(alive - {j}))
        expected.append(best)
        alive.remove(best)
    assert backward_elimination(torch.linalg.inv(h), q, 3).tolist() == expected


def test_backward_elimination_respects_the_limit_on_marked_unknowns():
    g = torch.Generator().manual_seed(5)
    x = torch.randn(30, 6, generator=g, dtype=torch.float64)
    h, q = x.T @ x, x.T @ torch.randn(30, generator=g, dtype=torch.float64)
    limited = torch.tensor([True, True, True, False, False, False])
    removed = backward_elimination(torch.linalg.inv(h), q, 4, limited, 1)
    assert int(limited[removed].sum()) <= 1 and len(set(removed.tolist())) == 4


def test_regularized_inverse_shifts_by_a_fraction_of_the_largest_eigenvalue():
    h = torch.diag(torch.tensor([0.0, 1.0, 2.0, 3.0, 4.0, 10.0], dtype=torch.float64))
    inverse = regularized_inverse(h, 1e-3)
    torch.testing.assert_close(inverse, torch.linalg.inv(h + 0.01 * torch.eye(6, dtype=torch.float64)), atol=1e-8, rtol=1e-6)


@pytest.mark.parametrize("seed", [1, 3, 4, 28])
def test_regularized_inverse_has_a_positive_scalar_shift_and_scales_with_the_gram(seed):
    g = torch.Generator().manual_seed(seed)
    x = torch.randn(30, 6, dtype=torch.float64, generator=g)
    h = x.T @ x
    inverse = regularized_inverse(h, 1e-3)
    difference = torch.linalg.inv(inverse) - h
    shift = difference.diag().mean()
    assert 0 < shift <= 1e-3 * torch.linalg.eigvalsh(h)[-1] * (1 + 1e-10)
    torch.testing.assert_close(difference, shift * torch.eye(6, dtype=h.dtype), atol=1e-12, rtol=1e-9)
    torch.testing.assert_close(regularized_inverse(7 * h, 1e-3) * 7, inverse)


def test_utility_marks_a_duplicated_column_redundant():
    x = torch.randn(40, 4, dtype=torch.float64, generator=torch.Generator().manual_seed(2))
    x = torch.cat((x, x[:, :1]), 1)
    efficiency = 2 * utility(x.T @ x, torch.zeros(5, dtype=torch.float64))
    assert efficiency[0] < 1e-6 and efficiency[4] < 1e-6 and (efficiency[1:4] > 0.1).all()


def test_pcg_solves_a_spd_system():
    g = torch.Generator().manual_seed(7)
    x = torch.randn(30, 10, dtype=torch.float64, generator=g)
    h = x.T @ x + 0.1 * torch.eye(10, dtype=torch.float64)
    b = torch.randn(10, dtype=torch.float64, generator=g)
    solution = pcg(lambda v: h @ v, b, h.diag(), iterations=200, tolerance=1e-12)
    torch.testing.assert_close(solution, torch.linalg.solve(h, b), atol=1e-9, rtol=1e-9)


def _pcg_early_exit(operator, rhs, diagonal, iterations, tolerance):
    """Conjugate gradients that leave the loop at either test, for comparison."""
    x, residual = torch.zeros_like(rhs), rhs.clone()
    goal = tolerance * rhs.norm()
    z = residual / diagonal
    direction, rz = z.clone(), residual @ z
    for _ in range(iterations):
        image = operator(direction)
        curvature = direction @ image
        if not curvature > 0:
            break
        step = rz / curvature
        x = x + step * direction
        residual = residual - step * image
        if residual.norm() <= goal:
            break
        z = residual / diagonal
        rz, previous = residual @ z, rz
        direction = z + (rz / previous) * direction
    return x


@pytest.mark.parametrize("case", ["converges", "indefinite", "limit", "nan_after_stop"])
def test_pcg_with_device_side_stops_equals_the_early_exit_loop(case):
    """The stopping tests run as masks on the device; the result is the early-exit loop's bit
    for bit, whether it stops on the tolerance, on non-positive curvature, at the iteration
    limit, or before an operator that would turn to NaN."""
    g = torch.Generator().manual_seed(11)
    x = torch.randn(40, 12, dtype=torch.float64, generator=g)
    h = x.T @ x + 0.05 * torch.eye(12, dtype=torch.float64)
    if case == "indefinite":
        h[5, 5] = -50.0
    b = torch.randn(12, dtype=torch.float64, generator=g)
    calls = []

    def operator(v):
        calls.append(1)
        out = h @ v
        return out * float("nan") if case == "nan_after_stop" and len(calls) > 3 else out

    iterations, tolerance = {"converges": (200, 1e-6), "indefinite": (50, 1e-12), "limit": (4, 1e-12),
                             "nan_after_stop": (50, 0.5)}[case]  # fmt: skip
    got = pcg(operator, b, h.diag().abs(), iterations=iterations, tolerance=tolerance)
    assert len(calls) == iterations
    calls.clear()
    want = _pcg_early_exit(operator, b, h.diag().abs(), iterations, tolerance)
    assert len(calls) < iterations or case == "limit"
    assert torch.equal(got, want)


def test_unit_vectors_with_a_frozen_component_stay_unit():
    spec = ParameterConfig(trainable=(True, True, False), normalize=True)
    reference = torch.nn.functional.normalize(torch.tensor([[0.3, 0.4, -0.6]]), dim=1)
    value = spec.project(torch.tensor([[2.0, -1.0, 5.0]]), reference)
    assert value[0, 2] == reference[0, 2]
    torch.testing.assert_close(value.norm(dim=1), torch.ones(1))


def test_unit_vector_tangent_matches_the_projection_with_a_frozen_component():
    spec = ParameterConfig(trainable=(True, False, True), normalize=True)
    value = torch.nn.functional.normalize(torch.randn(4, 3, generator=torch.Generator().manual_seed(3), dtype=torch.float64), dim=1)
    for axis in (0, 2):
        step = torch.zeros_like(value)
        step[:, axis] = 1e-6
        numeric = (spec.project(value + step, value) - spec.project(value - step, value)) / 2e-6
        torch.testing.assert_close(spec.tangent(step / 1e-6, value), numeric, atol=1e-7, rtol=0)


def test_a_free_part_of_zero_length_takes_the_reference_direction():
    spec = ParameterConfig(trainable=(True, False, True), normalize=True)
    reference = torch.tensor([[0.6, 0.8, 0.0]])
    torch.testing.assert_close(spec.project(torch.tensor([[0.0, 0.8, 0.0]]), reference), reference)


@pytest.mark.parametrize("trainable", [True, (True, False, True)])
def test_normalizing_without_any_reference_direction_raises(trainable):
    spec = ParameterConfig(trainable=trainable, normalize=True)
    with pytest.raises(ValueError, match="reference direction"):
        spec.project(torch.zeros(1, 3), torch.zeros(1, 3))


def test_bounds_period_and_frozen_components():
    spec = ParameterConfig(trainable=(True, False), lower=(-1.0, -5.0), upper=(1.0, 5.0))
    value = spec.project(torch.tensor([[3.0, 9.0], [-2.0, 0.0]]), torch.tensor([[0.0, 0.5], [0.0, -0.5]]))
    assert value.tolist() == [[1.0, 0.5], [-1.0, -0.5]]
    periodic = ParameterConfig(period=(10.0,))
    assert periodic.project(torch.tensor([12.5, -1.0]), torch.zeros(2)).tolist() == [2.5, 9.0]



@pytest.mark.parametrize("separation", [0.0, 0.3])
def test_distinct_matches_the_sequential_greedy(separation):
    """Coincident candidates (as coarse peaks beside unmoved refinements) keep only their
    best copy; the vectorized coincidence rule equals the one-by-one greedy."""
    from dsplat.replacement import _distinct

    g = torch.Generator().manual_seed(5)
    points = torch.rand(600, 3, generator=g)
    points[300:450] = points[:150]  # exact duplicates
    scores = torch.rand(600, generator=g)
    scores[400] = scores[100]  # a tie between copies
    occupied = torch.rand(50, 3, generator=g)
    occupied[0] = points[7]
    got = _distinct(scores, points, occupied, None, separation, 200)

    order = sorted(range(600), key=lambda i: (-float(scores[i]), i))
    kept = []
    for i in order:
        d_occ = (occupied - points[i]).norm(dim=1)
        if bool(((d_occ < separation) | (d_occ == 0)).any()):
            continue
        if any(float((points[j] - points[i]).norm()) < separation or torch.equal(points[j], points[i]) for j in kept):
            continue
        kept.append(i)
        if len(kept) == 200:
            break
    assert got.tolist() == kept
