"""Pins the oracle's tridiagonal solver and pressure solve: tridv against
scipy.linalg.solve_banded, press against a dense solve of the same Neumann
problem, and the projection against the divergence it must remove."""
import math

import numpy as np
import pytest
from scipy.linalg import solve_banded

import cf_oracle as O


def test_tridv_matches_banded_solver():
    rng = np.random.default_rng(1)
    n, m = 17, 5
    sub = rng.uniform(0.5, 1.5, (n, m))
    sup = rng.uniform(0.5, 1.5, (n, m))
    diag = -(np.abs(sub) + np.abs(sup) + rng.uniform(0.1, 1.0, (n, m)))
    r = rng.standard_normal((n, m))
    x = O.tridv(sub, diag, sup, r)
    for j in range(m):
        ab = np.zeros((3, n))
        ab[0, 1:] = sup[:-1, j]
        ab[1] = diag[:, j]
        ab[2, :-1] = sub[1:, j]
        assert np.allclose(x[:, j], solve_banded((1, 1), ab, r[:, j]), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("mode", [(1, 0), (0, 1), (3, 5), (7, -4)])
def test_press_solves_neumann_poisson(mode):
    """p'' - k^2 p = rhs in z with dp/dz = 0 at both walls (ghost p(0)=p(1),
    p(n+1)=p(n)), discretised on the u levels: the dense matrix is built from
    the finite-difference stencil, not from the tridiagonal rows."""
    nx, ny, nz = 16, 16, 12
    pr = O.Params(nx, ny, nz)
    mx, my = mode
    x = np.arange(nx) * pr.xl / nx
    y = np.arange(ny) * pr.yl / ny
    Y, X = np.meshgrid(y, x, indexing="ij")
    kx = mx * 2 * math.pi / pr.xl
    ky = my * 2 * math.pi / pr.yl
    rng = np.random.default_rng(4)
    prof = rng.standard_normal(nz)
    rhs = prof[:, None, None] * np.cos(kx * X + ky * Y)[None]
    p, _ = O.press(rhs, pr)
    # expected: dense solve of (D2 - k^2) q = prof with Neumann ghosts
    dz = pr.dz
    A = np.zeros((nz, nz))
    for i in range(nz):
        A[i, i] = -2.0 / dz ** 2 - (kx * kx + ky * ky)
        if i > 0:
            A[i, i - 1] = 1.0 / dz ** 2
        else:
            A[i, i] += 1.0 / dz ** 2          # p(0) = p(1)
        if i < nz - 1:
            A[i, i + 1] = 1.0 / dz ** 2
        else:
            A[i, i] += 1.0 / dz ** 2          # p(n+1) = p(n)
    q = np.linalg.solve(A, prof)
    expect = q[:, None, None] * np.cos(kx * X + ky * Y)[None]
    assert np.allclose(p, expect, rtol=1e-10, atol=1e-12)


def test_press_pins_mean_mode_to_zero():
    pr = O.Params(16, 16, 8)
    rhs = np.ones((8, 16, 16)) * np.arange(8)[:, None, None]
    p, _ = O.press(rhs, pr)
    assert np.allclose(p, 0.0)


def test_stage_projection_removes_divergence():
    pr = O.Params(16, 16, 12, dt=0.01)
    st = O.set_state(pr, *O.initial_fields(pr))
    assert np.abs(O.divergence(st, pr)).max() > 1e-2      # noise is not solenoidal
    for s in (1, 2, 3):
        O.stage(st, pr, s)
        assert np.abs(O.divergence(st, pr)).max() < 1e-11
