"""Float64 Cholesky factor and solve kernels against LAPACK (numpy) on the same matrices.

Bar: factor and solution within 1e-10 relative of LAPACK's float64 results for
well-conditioned systems; info equal to LAPACK's first failing leading minor.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def spd(n, seed, cond_shift=1.0):
    g = np.random.default_rng(seed)
    a = g.standard_normal((n, n))
    return a @ a.T / max(n, 1) + cond_shift * np.eye(n)


@pytest.mark.parametrize("n", [1, 5, 63, 64, 65, 130, 700, 2304])
def test_factor_matches_lapack_and_reads_only_the_lower_triangle(n):
    from dsplat.kernels.solver import cholesky_ex

    h = spd(n, n)
    noisy = h + np.triu(np.full((n, n), 7.0), 1)  # garbage above the diagonal
    factor, info = cholesky_ex(torch.tensor(noisy, device="cuda"))
    torch.cuda.synchronize()
    want = np.linalg.cholesky(h)
    got = factor.cpu().numpy()
    assert int(info) == 0
    assert np.array_equal(np.triu(got, 1), np.zeros((n, n)))
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"n={n} factor rel err {err:.2e}")
    assert err < 1e-12


@pytest.mark.parametrize("n,bad", [(3, 0), (70, 65), (200, 128), (200, 199)])
def test_info_is_the_first_non_positive_pivot(n, bad):
    from dsplat.kernels.solver import cholesky_ex

    d = np.ones(n)
    d[bad] = -1.0
    h = np.diag(d)
    _, info = cholesky_ex(torch.tensor(h, device="cuda"))
    ref = torch.linalg.cholesky_ex(torch.tensor(h))[1]
    assert int(info) == bad + 1 == int(ref)


@pytest.mark.parametrize("n,m", [(1, 1), (64, 1), (129, 3), (700, 9), (2304, 1), (2304, 2)])
def test_solve_matches_lapack(n, m):
    from dsplat.kernels.solver import cholesky_ex, cholesky_solve

    h = spd(n, 100 + n)
    g = np.random.default_rng(n)
    q = g.standard_normal((n, m))
    factor, info = cholesky_ex(torch.tensor(h, device="cuda"))
    x = cholesky_solve(torch.tensor(q, device="cuda"), factor).cpu().numpy()
    want = np.linalg.solve(h, q)
    err = np.linalg.norm(x - want) / np.linalg.norm(want)
    print(f"n={n} m={m} solve rel err {err:.2e}")
    assert int(info) == 0 and err < 1e-10


def test_varpro_solve_with_the_native_factor_agrees_with_torch():
    from dsplat.varpro import solve

    h = torch.tensor(spd(500, 3, 1e-3), device="cuda")
    q = torch.tensor(np.random.default_rng(4).standard_normal(500), device="cuda")
    keep = q.clone()
    x = solve(h, q)
    assert torch.equal(q, keep)  # the right-hand side is not overwritten
    want = torch.linalg.solve(h.cpu(), q.cpu())
    assert float((x.cpu() - want).norm() / want.norm()) < 1e-9


@pytest.mark.parametrize("n", [1, 63, 65, 700, 2304])
def test_float32_factor_matches_lapack(n):
    """The float32 form (the first attempt of the VarPro solve): within 1e-5 of LAPACK's
    float64 factor of the same float32 matrix, zeros above the diagonal, info 0."""
    from dsplat.kernels.solver import cholesky_ex

    h = spd(n, n + 1).astype(np.float32)
    noisy = h + np.triu(np.full((n, n), 7.0, dtype=np.float32), 1)
    factor, info = cholesky_ex(torch.tensor(noisy, device="cuda"))
    torch.cuda.synchronize()
    assert factor.dtype == torch.float32
    want = np.linalg.cholesky(h.astype(np.float64))
    got = factor.cpu().numpy().astype(np.float64)
    assert int(info) == 0
    assert np.array_equal(np.triu(got, 1), np.zeros((n, n)))
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"n={n} float32 factor rel err {err:.2e}")
    assert err < 1e-5


@pytest.mark.parametrize("n,bad", [(3, 0), (200, 128)])
def test_float32_info_is_the_first_non_positive_pivot(n, bad):
    """A leading minor made indefinite at row `bad` fails there in float32 as in LAPACK."""
    from dsplat.kernels.solver import cholesky_ex

    h = spd(n, 7 * n + bad)
    h[bad, bad] = -1.0  # leading minors before `bad` stay positive definite
    h = h.astype(np.float32)
    _, info = cholesky_ex(torch.tensor(h, device="cuda"))
    torch.cuda.synchronize()
    assert int(info) == bad + 1


def _power_on_host(h, steps):
    """The reference's power iteration (varpro.py:31-41) as a float64 tensor loop on the CPU."""
    vector = (torch.arange(len(h), dtype=torch.float64) * 0.6180339887) % 1 + 0.5
    top = vector.norm()
    for _ in range(steps):
        vector = h @ vector
        top = vector.norm()
        vector = vector / top.clamp(min=torch.finfo(top.dtype).tiny)
    return float(top)


@pytest.mark.parametrize("n", [1, 100, 2291])
def test_power_iteration_matches_the_host_loop(n):
    """One launch per step on the device against the tensor loop on the CPU: the same estimate
    to 1e-12 relative (only the summation order differs), and within the power method's own
    convergence of numpy's largest eigenvalue."""
    from dsplat.varpro import POWER_STEPS, largest_eigenvalue

    h = spd(n, 31 + n)
    got = float(largest_eigenvalue(torch.tensor(h, device="cuda")))
    want = _power_on_host(torch.tensor(h), POWER_STEPS)
    print(f"n={n} power estimate {got:.15e} host {want:.15e}")
    assert abs(got - want) <= 1e-12 * want
    assert got <= np.linalg.eigvalsh(h)[-1] * (1 + 1e-12)


def test_power_iteration_of_a_zero_matrix_is_zero():
    """A zero matrix gives zero products; the clamp keeps the normalization finite."""
    from dsplat.varpro import largest_eigenvalue

    assert float(largest_eigenvalue(torch.zeros(70, 70, dtype=torch.float64, device="cuda"))) == 0.0
