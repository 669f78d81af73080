"""Pins the oracle's composed stage (comp1 -> comprhs -> press -> comp2):
z-mirror equivariance of the channel, y-reflection equivariance, the
momentum budget (fluxes telescope to forcing minus wall stress), a closed-form
tendency of a single Fourier mode in uniform flow, and the RK3 invariants
(zero-mean w and rw, e >= emin)."""
import math

import numpy as np
import pytest

import cf_oracle as O


def base(nx=16, nz=10, dt=0.01, horizontally_uniform_e=True, seed=7):
    pr = O.Params(nx, nx, nz, dt=dt)
    u, v, w, e = O.initial_fields(pr, seed=seed)
    if horizontally_uniform_e:
        e = np.broadcast_to(1.6e-4 * (1 + 0.3 * np.cos(np.arange(nz)))[:, None, None], e.shape).copy()
    return pr, u, v, w, e


def mirror_input(pr, u, v, w, e):
    """z -> zl - z: u levels k <-> nnz+1-k, faces f <-> nnz-f, w -> -w."""
    n = pr.nnz
    um = u[::-1].copy(); vm = v[::-1].copy()
    wm = np.zeros_like(w); em = np.zeros_like(e)
    wm[: n - 1] = -w[: n - 1][::-1]          # faces 1..n-1 ; w(n)=0
    em[: n - 1] = e[: n - 1][::-1]
    em[n - 1] = em[n - 2]                    # e(nnz) ghost, reset by upper()
    return um, vm, wm, em


def test_z_mirror_equivariance_of_a_stage():
    pr, u, v, w, e = base()
    n = pr.nnz
    s1 = O.set_state(pr, u, v, w, e)
    s2 = O.set_state(pr, *mirror_input(pr, u, v, w, e))
    for st in (s1, s2):
        O.stage(st, pr, 1)
    I = lambda st, k: O.interior(st, pr, k)  # noqa: E731
    assert np.allclose(I(s2, "u"), I(s1, "u")[::-1], atol=1e-12)
    assert np.allclose(I(s2, "v"), I(s1, "v")[::-1], atol=1e-12)
    assert np.allclose(I(s2, "p"), I(s1, "p")[::-1], atol=1e-10)
    assert np.allclose(I(s2, "ru"), I(s1, "ru")[::-1], atol=1e-11)
    assert np.allclose(I(s2, "w")[: n - 1], -I(s1, "w")[: n - 1][::-1], atol=1e-12)
    assert np.allclose(I(s2, "rw")[: n - 1], -I(s1, "rw")[: n - 1][::-1], atol=1e-11)
    assert np.allclose(I(s2, "e")[: n - 1], I(s1, "e")[: n - 1][::-1], atol=1e-14)
    assert np.allclose(I(s2, "re")[: n - 1], I(s1, "re")[: n - 1][::-1], atol=1e-13)
    assert np.allclose(s2.divz, s1.divz[::-1], rtol=1e-9, atol=1e-20)


def test_y_reflection_equivariance():
    pr, u, v, w, e = base(horizontally_uniform_e=False)
    refl = lambda a: np.roll(a[:, ::-1, :], 1, axis=1)  # noqa: E731  y_j -> y_{-j}
    s1 = O.set_state(pr, u, v, w, e)
    s2 = O.set_state(pr, refl(u), -refl(v), refl(w), refl(e))
    for s in (1, 2):
        O.stage(s1, pr, s)
        O.stage(s2, pr, s)
    I = lambda st, k: O.interior(st, pr, k)  # noqa: E731
    for k, sg in (("u", 1), ("v", -1), ("w", 1), ("e", 1), ("p", 1), ("ru", 1), ("rv", -1)):
        assert np.allclose(I(s2, k), sg * refl(I(s1, k)), atol=1e-11), k


def test_momentum_budget_telescopes():
    """sum_k <ru(k)> = 0.016 nnz - (<tau_uw(nnz)> - <tau_uw(0)>)/dz: the
    advective z-fluxes vanish at the walls (w = 0) and the SGS z-fluxes leave
    only the two wall stresses (Schumann, lowerbc.f90:129 / upperbc.f90:127)."""
    pr, u, v, w, e = base(horizontally_uniform_e=False)
    st = O.set_state(pr, u, v, w, e)
    O.lower(st, pr); O.upper(st, pr)
    u1 = np.mean(O.lev(st.u, 1)); v1 = np.mean(O.lev(st.v, 1))
    un = np.mean(O.lev(st.u, pr.nnz)); vn = np.mean(O.lev(st.v, pr.nnz))
    O.comp1(st, pr, 1, False, 0.0)
    total = np.sum([np.mean(O.lev(st.ru, k)) for k in range(1, pr.nnz + 1)])
    vs_lo = max(math.hypot(u1, v1), 0.01); ut_lo = vs_lo * O.VK / pr.zody
    vs_hi = max(math.hypot(un, vn), 0.01); ut_hi = vs_hi * O.VK / pr.zody
    tau0 = -ut_lo ** 2 / vs_lo * u1
    taun = ut_hi ** 2 / vs_hi * un
    expect = 0.016 * pr.nnz - (taun - tau0) * pr.dzm
    assert abs(total - expect) < 1e-12 * max(1.0, abs(expect))


def test_single_mode_tendency_closed_form():
    """u = U0 + a cos(kx x), v = w = 0, e uniform: interior levels have
    ru = 0.016 - d(u^2)/dx - d(tau_uu)/dx with tau_uu = -2 K du/dx,
    K = min(0.1 dsl sqrt(e), xkmax).  The a^2 harmonic at 2m survives the
    3/2-rule truncation only while 2m < nx/2."""
    nx, nz = 32, 8
    for m in (3, 11):
        pr = O.Params(nx, nx, nz, dt=0.01)
        x = np.arange(nx) * pr.xl / nx
        k = m * 2 * math.pi / pr.xl
        U0, a, e0 = 1.3, 0.05, 2.0e-4
        u = np.broadcast_to(U0 + a * np.cos(k * x)[None, None, :], (nz, nx, nx)).copy()
        z = np.zeros_like(u)
        st = O.set_state(pr, u, z, z, np.full_like(u, e0))
        O.lower(st, pr); O.upper(st, pr)
        O.comp1(st, pr, 1, False, 0.0)
        K = min(0.1 * pr.dsl * math.sqrt(e0), pr.xkmax)
        harm = a * a * k * np.sin(2 * k * x) if 2 * m < nx // 2 else 0.0
        expect = 0.016 + 2 * U0 * a * k * np.sin(k * x) + harm - 2 * K * a * k * k * np.cos(k * x)
        for lev in range(2, nz):
            got = O.lev(st.ru, lev)
            assert np.allclose(got, np.broadcast_to(expect, got.shape), atol=1e-12), (m, lev)
        assert np.allclose(O.interior(st, pr, "rv"), 0.0, atol=1e-13)
        # the SGS TKE: uniform e, no diffusion in x/y, only advection by the
        # horizontal flow and dissipation; e u advects as -0.5 d/dx(e (u + u_up))
        re = O.lev(st.re, 4)
        exp_re = -0.5 * e0 * 2 * (-a * k * np.sin(k * x)) - pr.ceps * e0 ** 1.5
        assert np.allclose(re, np.broadcast_to(exp_re, re.shape), atol=1e-14), m


def test_rk3_invariants():
    pr, u, v, w, e = base(horizontally_uniform_e=False)
    st = O.set_state(pr, u, v, w, e)
    for it in range(3):
        O.step(st, pr)
        n = pr.nnz
        assert np.abs(O.interior(st, pr, "w")[: n - 1].mean(axis=(1, 2))).max() < 1e-16
        assert np.abs(O.interior(st, pr, "rw")[: n - 1].mean(axis=(1, 2))).max() < 1e-15
        assert O.interior(st, pr, "e")[: n - 1].min() >= O.EMIN
        assert np.abs(O.lev(st.w, n)).max() == 0.0


def test_rk3_is_third_order_on_linear_decay():
    """The low-storage scheme (les.F90:26-27, comp1.f90:126, comp2.f90:34)
    applied to dphi/dt = lam phi has error O(dt^4) per step."""
    lam = -1.7
    errs = []
    for nsteps in (10, 20, 40):
        dt = 1.0 / nsteps
        phi, r = 1.0, 0.0
        for _ in range(nsteps):
            for s in range(3):
                rn = lam * phi                   # comp1 sees the stage-start field
                phi = phi - dt * O.ZETAS[s] * r + dt * O.GAMA[s] * rn
                r = rn
        errs.append(abs(phi - math.exp(lam)))
    order = math.log2(errs[0] / errs[1]), math.log2(errs[1] / errs[2])
    assert min(order) > 2.8, order
