"""The oracle's discrete adjoint (the checker for the GPU VJP): every
transpose satisfies the dot-product identity <A x, y> = <x, A^T y> under the
real inner product over stored entries, and the step VJP matches central
finite differences (the step is a polynomial in u).  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O


def ip(a, b):
    return float(np.vdot(a, b).real)


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("shape", [(8, 8, 8), (12, 8, 16), (6, 12, 10)])
def test_c2r_transpose_identity(shape):
    rng = np.random.default_rng(1)
    X = crandn(rng, (shape[0], shape[1], shape[2] // 2 + 1))  # non-Hermitian planes included
    g = rng.standard_normal(shape)
    lhs = ip(O.c2r_3d(X, shape[2]), g)
    rhs = ip(X, O.c2r_3d_T(g))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


@pytest.mark.parametrize("shape", [(8, 8, 8), (12, 8, 16), (6, 12, 10)])
def test_r2c_transpose_identity(shape):
    rng = np.random.default_rng(2)
    x = rng.standard_normal(shape)
    Y = crandn(rng, (shape[0], shape[1], shape[2] // 2 + 1))
    lhs = ip(O.r2c_3d(x), Y)
    rhs = ip(x, O.r2c_3d_T(Y, shape[2]))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


@pytest.mark.parametrize("n", [8, 16])
def test_nonlinear_vjp_matches_exact_jvp(n):
    """N(u) is quadratic, so the central difference (N(u+e d) - N(u-e d))/(2e)
    is its exact directional derivative; <JVP, lam> = <d, VJP>."""
    rng = np.random.default_rng(n)
    s = (3, n, n, n // 2 + 1)
    u, d, lam = crandn(rng, s), crandn(rng, s), crandn(rng, s)
    eps = 1e-3
    jvp = (O.nonlinear(u + eps * d) - O.nonlinear(u - eps * d)) / (2 * eps)
    lhs = ip(jvp, lam)
    rhs = ip(d, O.nonlinear_vjp(u, lam))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_step_vjp_matches_central_differences():
    n, nu, dt = 8, 0.02, 5e-3
    rng = np.random.default_rng(4)
    s = (3, n, n, n // 2 + 1)
    u = 0.05 * O.leray(crandn(rng, s)) * n ** 3
    d = O.leray(crandn(rng, s))
    lam = crandn(rng, s)

    def step(v):
        sol = O.Solver(n, nu)
        sol.set_state(v)
        sol.set_dt(dt, True)
        sol.step(1)
        return sol.state()

    sol = O.Solver(n, nu)
    sol.set_state(u)
    sol.set_dt(dt, True)
    g = sol.step_vjp(lam)
    eps = 1e-6 * np.linalg.norm(u) / np.linalg.norm(d)
    fd = ip((step(u + eps * d) - step(u - eps * d)) / (2 * eps), lam)
    assert abs(fd - ip(d, g)) < 1e-8 * abs(fd)
    assert np.array_equal(sol.state(), u)  # the VJP does not advance the state
