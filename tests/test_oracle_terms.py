"""Closed-form pins of the oracle's comp1 terms that the symmetry and budget
tests leave open: the w advection and the TKE fluxes of compadv1
(compadv1.f90:45-51, :59-63), the TKE diffusion and the w-equation stress
divergence of compsgs3 (compsgs3.f90:23-30, :56-59), the stale ex/ey at face
nnz-1 (derivs2.f90:9-21: ex, ey are not recomputed once iz >= nnz-1), and
the u/v-equation z fluxes with the wall stresses at both walls
(compadv1.f90:22-24, compsgs3.f90:40-42, compsgs1.f90:41-50).

Each test builds fields that are low trigonometric polynomials in x (or y)
with level-dependent amplitudes, so that every product stays inside the
3/2-grid band and compress (filter_fftw.f90:285-308) is an exact truncation.
The expected tendency is then written out by hand from the product-to-sum
identities, with the derivatives taken analytically: no FFT and no oracle
function is used on the expected side.  The amplitudes vary with the level
so that a staggering slip (u(k-1) for u(k+1), face k for face k-1, K at the
wrong face) changes the answer, and the factors 1/2 and 1/(2 dz) appear with
distinct coefficients so that a dropped or doubled factor does too.

Mutations of the oracle that these tests catch (each planted by hand in a
copy of oracle/cf_oracle.py on the CPU and seen to fail):
  * -0.5 d/dx((u+u_up) w)  ->  -1.0 d/dx(...)           test_w_and_tke_fluxes
  * (u0 + up) * w0         ->  (u0 + um) * w0              test_w_and_tke_fluxes
  * (wp*ep - wm*em)*todzm  ->  (wp*ep - wm*em)*dzm         test_w_and_tke_fluxes
  * (wp*ep - wm*em)        ->  (wp*e0 - wm*e0)             test_w_and_tke_fluxes
  * xderiv(2 K ex)         ->  xderiv(K ex)                test_tke_diffusion_and_stale_gradient
  * (xkp+xk0)(ep-e0)       ->  (xkm+xk0)(ep-e0)            test_tke_diffusion_and_stale_gradient
  * ex(nnz-1) recomputed instead of stale                   test_tke_diffusion_and_stale_gradient
  * tauww[k+1]-tauww[k]    ->  tauww[k]-tauww[k-1]         test_w_and_tke_fluxes
  * tauuw[k0] in rw        ->  tauuw[km]                   test_w_and_tke_fluxes
  * (w-bar^2 difference)*dzm  ->  *todzm                   test_w_and_tke_fluxes
  * e0 (u0 + up) in re     ->  e0 (u0 + um)                both tests
"""
import math

import numpy as np
import pytest

import cf_oracle as O

NX, NZ = 16, 8


def _grid(pr):
    x = np.arange(pr.nnx) * pr.xl / pr.nnx
    y = np.arange(pr.nny) * pr.yl / pr.nny
    return x[None, :], y[:, None]                       # broadcast to (y, x)


def _direction(pr, d):
    """cos, sin and the second harmonic of the lowest mode along x or y, and
    its physical wavenumber (y has length pi, so its lowest mode is cos 2y)."""
    x, y = _grid(pr)
    if d == "x":
        kap = 2 * math.pi / pr.xl
        s = kap * x + 0 * y
    else:
        kap = 2 * math.pi / pr.yl
        s = kap * y + 0 * x
    return np.cos(s), np.sin(s), np.cos(2 * s), np.sin(2 * s), kap


def _run_comp1(pr, u, v, w, e):
    st = O.set_state(pr, u, v, w, e)
    O.lower(st, pr)
    O.upper(st, pr)
    O.comp1(st, pr, 1, False, 0.0)
    return st


def _ext(a, lo_val, hi_val, n):
    """Per-level amplitudes a[1..n] extended to indices 0..n+1."""
    out = np.zeros(n + 2)
    out[1: n + 1] = a
    out[0] = lo_val
    out[n + 1] = hi_val
    return out


@pytest.mark.parametrize("d", ["x", "y"])
@pytest.mark.parametrize("alk", [0.0, None])
def test_w_and_tke_fluxes(d, alk):
    """Along-flow velocity q_k = Q_k + a_k T (q = u for d = x, v for d = y),
    w_k = b_k T at faces 1..nnz-1, e_k = E_k uniform per face, imodel 3 with
    K_k = 0.1 alk sqrt(E_k) (alk = 0: advection alone).  T = cos(kap s).

    rw_k = 0.5 kap S_k b_k sinT + 0.5 kap A_k b_k sin2T
           - dz^-1/8 [(b_k+1 + b_k)^2 - (b_k + b_k-1)^2] cos2T         (w-bar^2 flux)
           - kap K_k (a_k+1 - a_k)/dz sinT - kap^2 K_k b_k T            (d tau_qw / ds)
           + dz^-2 [(K_k + K_k+1)(b_k+1 - b_k) - (K_k-1 + K_k)(b_k - b_k-1)] T
    re_k = 0.5 kap E_k A_k sinT - (b_k+1 E_k+1 - b_k-1 E_k-1)/(2 dz) T
           + dz^-2 [(K_k+1 + K_k)(E_k+1 - E_k) - (K_k + K_k-1)(E_k - E_k-1)] - ceps E_k^1.5
    with S_k = Q_k + Q_k+1, A_k = a_k + a_k+1, and the along-flow tendency
    rq_k = 0.016[x] - d(q^2)/ds - z-flux of q w - d tau_qq/ds - (tau_qw,k - tau_qw,k-1)/dz
    including the Schumann wall stresses of faces 0 and nnz
    (tau = -/+ utau^2/vsfc q_wall, lowerbc.f90:128-130 and its upper twin)."""
    pr = O.Params(NX, NX, NZ, dt=0.01, alk=alk)
    n, dzm, todzm = pr.nnz, pr.dzm, pr.todzm
    T, S, T2, S2, kap = _direction(pr, d)
    k = np.arange(1, n + 1)
    Q = 0.6 + 0.13 * k + 0.011 * k * k                   # level means of q
    a = 0.04 * (1.0 + 0.35 * np.sin(1.3 * k))            # level amplitudes of q
    b = 0.03 * (1.0 + 0.5 * np.cos(0.9 * k))             # face amplitudes of w
    b[n - 1] = 0.0                                       # w(nnz) = 0: the top wall
    E = 2.0e-4 * (1.0 + 0.4 * np.sin(0.7 * k))           # face values of e
    q = Q[:, None, None] + a[:, None, None] * T[None]
    zero = np.zeros((n, pr.nny, pr.nnx))
    wf = b[:, None, None] * T[None] + zero
    ef = E[:, None, None] + zero
    u, v = (q, zero) if d == "x" else (zero, q)
    st = _run_comp1(pr, u, v, wf, ef)

    # extended per-index profiles: levels 0..nnz+1 (ghosts u,v = 0 below and above)
    Qx = _ext(Q, 0.0, 0.0, n)
    ax = _ext(a, 0.0, 0.0, n)
    bx = _ext(b, 0.0, 0.0, n)                            # faces 0..nnz+1, w = 0 at the walls
    bx[n] = 0.0
    Ex = _ext(E, E[0], E[n - 2], n)                      # e(0) = e(1), e(nnz) = e(nnz-1)
    Ex[n] = E[n - 2]
    Kx = 0.1 * pr.alk * np.sqrt(Ex)                      # faces 0..nnz
    Kx = np.minimum(Kx, pr.xkmax)

    for kk in range(1, n):                               # faces 1..nnz-1
        Sk = Qx[kk] + Qx[kk + 1]
        Ak = ax[kk] + ax[kk + 1]
        rw = (0.5 * kap * Sk * bx[kk] * S + 0.5 * kap * Ak * bx[kk] * S2
              - 0.125 * dzm * ((bx[kk + 1] + bx[kk]) ** 2 - (bx[kk] + bx[kk - 1]) ** 2) * T2
              - kap * Kx[kk] * (ax[kk + 1] - ax[kk]) * dzm * S - kap * kap * Kx[kk] * bx[kk] * T
              + dzm * dzm * ((Kx[kk] + Kx[kk + 1]) * (bx[kk + 1] - bx[kk])
                             - (Kx[kk - 1] + Kx[kk]) * (bx[kk] - bx[kk - 1])) * T)
        got = O.lev(st.rw, kk)
        assert np.allclose(got, rw + 0 * got, rtol=0, atol=1e-13), ("rw", d, alk, kk)
        re = (0.5 * kap * Ex[kk] * Ak * S
              - (bx[kk + 1] * Ex[kk + 1] - bx[kk - 1] * Ex[kk - 1]) * todzm * T
              + ((Kx[kk + 1] + Kx[kk]) * (Ex[kk + 1] - Ex[kk])
                 - (Kx[kk] + Kx[kk - 1]) * (Ex[kk] - Ex[kk - 1])) * dzm * dzm
              - pr.ceps * Ex[kk] ** 1.5)
        got = O.lev(st.re, kk)
        assert np.allclose(got, re + 0 * got, rtol=0, atol=1e-15), ("re", d, alk, kk)

    # along-flow momentum, levels 1..nnz, with the wall stresses at faces 0 and nnz
    vs_lo = max(abs(Q[0]), 0.01)
    ut_lo = vs_lo * O.VK / pr.zody
    vs_hi = max(abs(Q[n - 1]), 0.01)
    ut_hi = vs_hi * O.VK / pr.zody

    def tau_qw(f):                                       # face f, as (const, cos, sin) coefficients
        if f == 0:
            c = -ut_lo ** 2 / vs_lo
            return c * Qx[1], c * ax[1], 0.0
        if f == n:
            c = ut_hi ** 2 / vs_hi
            return c * Qx[n], c * ax[n], 0.0
        K = Kx[f]
        return (-K * (Qx[f + 1] - Qx[f]) * dzm, -K * (ax[f + 1] - ax[f]) * dzm, K * kap * bx[f])

    for kk in range(1, n + 1):
        Kc = Kx[kk - 1] + Kx[kk]
        rq = (0.016 if d == "x" else 0.0) + 2 * kap * Qx[kk] * ax[kk] * S + kap * ax[kk] ** 2 * S2
        up_c = (Qx[kk + 1] + Qx[kk]) * bx[kk]            # (q_k+1 + q_k) w_k = up_c T + up_h (1 + cos2T)
        up_h = 0.5 * (ax[kk + 1] + ax[kk]) * bx[kk]
        dn_c = (Qx[kk - 1] + Qx[kk]) * bx[kk - 1]
        dn_h = 0.5 * (ax[kk - 1] + ax[kk]) * bx[kk - 1]
        rq = rq - ((up_c - dn_c) * T + (up_h - dn_h) * (1.0 + T2)) * todzm
        rq = rq - Kc * kap * kap * ax[kk] * T            # tau_qq = Kc a kap sinT
        c1, cc1, s1 = tau_qw(kk)
        c0, cc0, s0 = tau_qw(kk - 1)
        rq = rq - ((c1 - c0) + (cc1 - cc0) * T + (s1 - s0) * S) * dzm
        got = O.lev(st.ru if d == "x" else st.rv, kk)
        assert np.allclose(got, rq + 0 * got, rtol=0, atol=1e-12), ("rq", d, alk, kk)
        other = O.lev(st.rv if d == "x" else st.ru, kk)
        assert np.allclose(other, (0.0 if d == "x" else 0.016) + 0 * other, atol=1e-13)


@pytest.mark.parametrize("d", ["x", "y"])
@pytest.mark.parametrize("profile", ["all", "only_top", "only_below_top"])
def test_tke_diffusion_and_stale_gradient(d, profile):
    """imodel 4 with u = U_k uniform per level: the face shear is uz_k^2 alone,
    so K_k = min((0.1 dsl)^2 |U_k+1 - U_k|/dz, xkmax) varies with the face and
    not in x or y (complen.f90:13-23; uz(0) = uz(1), uz(nnz) = uz(nnz-1),
    derivs1.f90:37, derivs2.f90:27).  e_k = E_k + eps_k T, no dissipation
    (ceps = 0), v = w = 0:

    re_k = 0.5 kap eps_k (U_k + U_k+1) sinT [d = x]
           - 2 kap^2 K_k eps'_k T
           + dz^-2 [(K_k+1 + K_k)(e_k+1 - e_k) - (K_k + K_k-1)(e_k - e_k-1)]

    where eps'_k = eps_k, except at face nnz-1, where ex and ey are the ones of
    face nnz-2 (derivs2.f90:9-21 computes them only for iz < nnz-1):
    eps'_nnz-1 = eps_nnz-2.  The 'only_top' profile puts the horizontal
    variation of e at face nnz-1 alone (its horizontal diffusion must vanish);
    'only_below_top' puts it at face nnz-2 alone (face nnz-1 must still see it)."""
    pr = O.Params(NX, NX, NZ, dt=0.01, imodel=4)
    pr.ceps = 0.0
    n, dzm = pr.nnz, pr.dzm
    T, S, T2, S2, kap = _direction(pr, d)
    k = np.arange(1, n + 1)
    U = 0.4 + 0.21 * k + 0.017 * k * k
    E = 3.0e-4 * (1.0 + 0.3 * np.cos(1.1 * k))
    eps = {"all": 4.0e-5 * (1.0 + 0.5 * np.sin(0.8 * k)),
           "only_top": np.where(k == n - 1, 5.0e-5, 0.0),
           "only_below_top": np.where(k == n - 2, 5.0e-5, 0.0)}[profile]
    zero = np.zeros((n, pr.nny, pr.nnx))
    u = U[:, None, None] + zero
    e = E[:, None, None] + eps[:, None, None] * T[None]
    st = _run_comp1(pr, u, zero, zero, e)

    Ux = _ext(U, 0.0, 0.0, n)
    Ex = _ext(E, E[0], E[n - 2], n)
    Ex[n] = E[n - 2]
    ex = _ext(eps, eps[0], eps[n - 2], n)
    ex[n] = eps[n - 2]
    uz = np.zeros(n + 1)                                  # faces 0..nnz
    uz[1:n] = (Ux[2:n + 1] - Ux[1:n]) * dzm
    uz[0] = uz[1]
    uz[n] = uz[n - 1]
    K = np.minimum((0.1 * pr.dsl) ** 2 * np.abs(uz), pr.xkmax)
    for kk in range(1, n):
        eprime = ex[kk] if kk < n - 1 else ex[n - 2]
        re = (-2.0 * kap * kap * K[kk] * eprime * T
              + ((K[kk + 1] + K[kk]) * (Ex[kk + 1] - Ex[kk]) - (K[kk] + K[kk - 1]) * (Ex[kk] - Ex[kk - 1])) * dzm * dzm
              + ((K[kk + 1] + K[kk]) * (ex[kk + 1] - ex[kk]) - (K[kk] + K[kk - 1]) * (ex[kk] - ex[kk - 1])) * dzm * dzm * T)
        if d == "x":
            re = re + 0.5 * kap * ex[kk] * (Ux[kk] + Ux[kk + 1]) * S
        got = O.lev(st.re, kk)
        assert np.allclose(got, re + 0 * got, rtol=0, atol=1e-16), (d, profile, kk)
