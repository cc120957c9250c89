"""The host flexible GMRES of the coupled adjoint (exfemus_amd.coupled.fgmres) on small dense systems: a
nonsymmetric matrix with a Jacobi preconditioner, and a preconditioner that changes from call to call
(truncated inner iterations, as the multigrid block solves are), against numpy's direct solve."""
import numpy as np

from exfemus_amd import coupled


def _system(n=200, seed=3):
    rng = np.random.default_rng(seed)
    A = np.eye(n) * 4.0 + rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    return A, b


def test_fgmres_jacobi():
    A, b = _system()
    d = np.diag(A)
    x, its, rel = coupled.fgmres(lambda v: A @ v, lambda r: r / d, b, tol=1e-12, maxit=200)
    assert rel <= 1e-12
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-11)
    assert its < 60


def test_fgmres_variable_preconditioner_and_restart():
    A, b = _system(n=300, seed=7)
    d = np.diag(A)
    count = [0]

    def precond(r):          # 1 to 3 Jacobi sweeps, varying: a flexible method is needed
        count[0] += 1
        z = r / d
        for _ in range(count[0] % 3):
            z = z + (r - A @ z) / d
        return z

    x, its, rel = coupled.fgmres(lambda v: A @ v, precond, b, tol=1e-11, maxit=300, restart=7)
    assert rel <= 1e-11
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-8, atol=1e-10)
