"""Closed-form pins of the oracle's diagnostics: compmn (compmn.F90:17-62),
the SGS stress means of an mnout stage (compsgs3.f90:71-78) and the x-y
spectra of calc_spectra (calc_spectra.f90:139-257).

The expected values are derived by hand from constructed fields (plane means
of products of trigonometric polynomials, Parseval, and ring counts found by
brute-force enumeration of the integer wavenumber plane); no oracle function
computes them.

Mutations of the oracle these tests catch (planted by hand in a copy of
oracle/cf_oracle.py on the CPU and seen to fail):
  * compmn: ul from u(k-1) instead of u(k+1)                 test_compmn_constructed
  * compmn: wwle from the face average of w, not w(k)        test_compmn_constructed
  * compmn: the 0.5 of englez dropped                        test_compmn_constructed
  * SGS means: uwsb from face k-1 instead of face k          test_sgs_means_of_an_mnout_stage
  * calc_spectra: weight 2 on the kx Nyquist column          test_spectra_parseval / single_mode
  * calc_spectra: u averaged over zl-1, zl                   test_spectra_single_mode
  * calc_spectra: the pooled nCount not divided by 3 fields  test_spectra_single_mode
"""
import math

import numpy as np
import pytest

import cf_oracle as O


def _xy(nx, ny, xl, yl):
    x = (np.arange(nx) * xl / nx)[None, :]
    y = (np.arange(ny) * yl / ny)[:, None]
    return x, y


def test_compmn_constructed():
    """u_k = U_k + a_k cos x + c_k sin 2y, v_k = V_k + d_k cos(x+2y) + h_k sin 2y,
    w_k = W_k + b_k cos x + g_k cos(x+2y), e_k = E_k + eps_k cos x.
    With ul_k = ((u_k - <u_k>) + (u_k+1 - <u_k+1>))/2 (u(nnz+1) = 0):
      uule = ((a_k+a_k+1)^2 + (c_k+c_k+1)^2)/8,  vvle = ((d_k+d_k+1)^2 + (h_k+h_k+1)^2)/8,
      wwle = W_k^2 + (b_k^2 + g_k^2)/2   (w at face k itself, not its face average),
      uwle = (a_k+a_k+1) b_k/4, vwle = (d_k+d_k+1) g_k/4, uvle = (c_k+c_k+1)(h_k+h_k+1)/8,
      englez = (uule + vvle + wwle)/2, and the plane means U, V, W, E."""
    pr = O.Params(16, 16, 6)
    n = pr.nnz
    x, y = _xy(pr.nnx, pr.nny, pr.xl, pr.yl)
    k = np.arange(1, n + 1)
    U = 1.0 + 0.2 * k; V = 0.1 - 0.03 * k; W = 0.01 * np.sin(k); E = 1e-3 * (1 + 0.1 * k)
    a = 0.05 + 0.01 * k; c = 0.02 * np.cos(k); d = 0.03 + 0.002 * k * k; h = 0.04 * np.sin(2 * k)
    b = 0.02 + 0.005 * k; g = 0.01 * (1 + np.cos(k)); eps = 1e-4 * k
    W[n - 1] = b[n - 1] = g[n - 1] = 0.0
    L = lambda A: A[:, None, None]  # noqa: E731
    u = L(U) + L(a) * np.cos(x) + L(c) * np.sin(2 * y)
    v = L(V) + L(d) * np.cos(x + 2 * y) + L(h) * np.sin(2 * y)
    w = L(W) + L(b) * np.cos(x) + L(g) * np.cos(x + 2 * y)
    e = L(E) + L(eps) * np.cos(x) + 0 * y
    st = O.set_state(pr, u, v, w, e)
    st.divz = np.linspace(1e-30, 2e-30, n)
    st.sgs = {key: np.full(n, 0.1 * (i + 1)) for i, key in enumerate(("uusb", "vvsb", "uvsb", "wwsb", "uwsb", "vwsb"))}
    mb = O.compmn(st, pr)
    nx = lambda A: np.append(A[1:], 0.0)  # noqa: E731  amplitude at k+1, 0 above the top
    uule = ((a + nx(a)) ** 2 + (c + nx(c)) ** 2) / 8
    vvle = ((d + nx(d)) ** 2 + (h + nx(h)) ** 2) / 8
    wwle = W ** 2 + (b ** 2 + g ** 2) / 2
    expect = dict(uxym=U, vxym=V, wxym=W, engsbz=E, uule=uule, vvle=vvle, wwle=wwle,
                  uwle=(a + nx(a)) * b / 4, vwle=(d + nx(d)) * g / 4,
                  uvle=(c + nx(c)) * (h + nx(h)) / 8, englez=(uule + vvle + wwle) / 2,
                  divz=st.divz, pple=np.zeros(n))
    for i, key in enumerate(("uusb", "vvsb", "uvsb", "wwsb", "uwsb", "vwsb")):
        expect[key] = np.full(n, 0.1 * (i + 1))
    for row, name in enumerate(O.MEAN_NAMES):
        assert np.allclose(mb[row], expect[name], rtol=1e-12, atol=1e-16), name


def test_sgs_means_of_an_mnout_stage():
    """imodel 3, e_k = E_k + eps_k sin x, u_k = U_k + a_k cos x + s_k sin x,
    v = 0, w_k = b_k cos x + p_k sin x, so the 3/2-grid expanded fields equal
    these analytic functions at the 3/2-grid points.  The stress means are then
    plane means of pointwise products there (sum * xymb, compsgs3.f90:71-78):
    uusb_k = <-(K_k-1 + K_k) ux_k>, wwsb_k = <-(K_k-1 + K_k) wz_k>,
    uwsb_k = <tau_uw at face k> (the Schumann stress at face nnz), uvsb = vvsb =
    vwsb = 0, with K_f = min(0.1 dsl sqrt(e_f), xkmax) (complen.f90:21)."""
    pr = O.Params(16, 16, 6, dt=0.01)
    n = pr.nnz
    x, y = _xy(pr.nnx, pr.nny, pr.xl, pr.yl)
    xb, yb = _xy(pr.nnxb, pr.nnyb, pr.xl, pr.yl)
    k = np.arange(1, n + 1)
    U = 0.9 + 0.1 * k; a = 0.05 * (1 + 0.2 * k); s = 0.03 * np.cos(k)
    b = 0.02 * (1 + 0.3 * np.sin(k)); p = 0.01 * k; E = 3e-4 * (1 + 0.2 * np.cos(k)); eps = 1e-4 * (1 + 0.1 * k)
    b[n - 1] = p[n - 1] = 0.0
    L = lambda A: A[:, None, None]  # noqa: E731
    zero = np.zeros((n, pr.nny, pr.nnx))
    u = L(U) + L(a) * np.cos(x) + L(s) * np.sin(x) + 0 * y
    w = L(b) * np.cos(x) + L(p) * np.sin(x) + 0 * y
    e = L(E) + L(eps) * np.sin(x) + 0 * y
    st = O.set_state(pr, u, zero, w, e)
    O.lower(st, pr); O.upper(st, pr)
    O.comp1(st, pr, 1, True, 0.0)

    def ef(f):                                          # e on the 3/2 grid at face f (ghosts e(0)=e(1), e(nnz)=e(nnz-1))
        f = min(max(f, 1), n - 1)
        return np.maximum(E[f - 1] + eps[f - 1] * np.sin(xb) + 0 * yb, O.EMIN)

    K = lambda f: np.minimum(0.1 * pr.dsl * np.sqrt(ef(f)), pr.xkmax)  # noqa: E731
    ub = lambda L_: (U[L_ - 1] + a[L_ - 1] * np.cos(xb) + s[L_ - 1] * np.sin(xb) + 0 * yb) if 1 <= L_ <= n else 0 * xb + 0 * yb  # noqa: E731
    wb = lambda f: (b[f - 1] * np.cos(xb) + p[f - 1] * np.sin(xb) + 0 * yb) if 1 <= f <= n - 1 else 0 * xb + 0 * yb  # noqa: E731
    mean = lambda A: A.sum() * pr.xymb  # noqa: E731
    for L_ in range(1, n + 1):
        ux = -a[L_ - 1] * np.sin(xb) + s[L_ - 1] * np.cos(xb) + 0 * yb
        wz = (wb(L_) - wb(L_ - 1)) * pr.dzm
        Kc = K(L_ - 1) + K(L_)
        assert math.isclose(st.sgs["uusb"][L_ - 1], mean(-Kc * ux), rel_tol=1e-11, abs_tol=1e-20), L_
        assert math.isclose(st.sgs["wwsb"][L_ - 1], mean(-Kc * wz), rel_tol=1e-11, abs_tol=1e-20), L_
        if L_ < n:
            uz = (ub(L_ + 1) - ub(L_)) * pr.dzm
            wx = -b[L_ - 1] * np.sin(xb) + p[L_ - 1] * np.cos(xb)
            tuw = mean(-K(L_) * (uz + wx))
        else:
            vs = max(abs(U[n - 1]), 0.01)
            ut = vs * O.VK / pr.zody
            tuw = ut * ut / vs * U[n - 1]
        assert math.isclose(st.sgs["uwsb"][L_ - 1], tuw, rel_tol=1e-11, abs_tol=1e-20), L_
        for key in ("vvsb", "uvsb", "vwsb"):
            assert abs(st.sgs[key][L_ - 1]) < 1e-20, key


def _ring_counts(N):
    """Weighted ring populations of one r2c transform: every (kx, ky) of the full
    plane kx, ky in [-N/2+1, N/2], binned by nint(sqrt(kx^2+ky^2)) <= N/2."""
    C = np.zeros(N // 2 + 1)
    for kx in range(-N // 2 + 1, N // 2 + 1):
        for ky in range(-N // 2 + 1, N // 2 + 1):
            r = int(math.floor(math.sqrt(kx * kx + ky * ky) + 0.5))
            if r <= N // 2:
                C[r] += 1
    return C


def _fields(N, nz, rng):
    return [rng.standard_normal((nz, N, N)) for _ in range(3)]


def test_spectra_parseval():
    """var = mean of the transformed plane squared (Parseval with the r2c
    weights 1, 2, ..., 2, 1 over kx = 0..N/2), for u and v the mean-removed
    average of levels zl and zl+1, for w the face zl itself."""
    N, nz = 16, 6
    rng = np.random.default_rng(4)
    u, v, w = _fields(N, nz, rng)
    zl = [1, 3, 5]
    hats, var, ls = O.calc_spectra([(u, v, w)], zl)
    for iz, z in enumerate(zl):
        pu = 0.5 * ((u[z - 1] - u[z - 1].mean()) + (u[z] - u[z].mean()))
        pv = 0.5 * ((v[z - 1] - v[z - 1].mean()) + (v[z] - v[z].mean()))
        for f, plane in enumerate((pu, pv, w[z - 1])):
            assert math.isclose(var[f][iz], np.mean(plane ** 2), rel_tol=1e-12), (f, z)


def test_spectra_ring_sums_against_the_full_plane():
    """The reference bins the r2c half plane with weights 1 (kx = 0, N/2) and 2
    (kx = 1..N/2-1).  The same ring sums follow from the full complex plane,
    every coefficient counted once: raw(r) = sum |F(kx,ky)|^2 over
    nint(sqrt(kx^2+ky^2)) = r <= N/2.  Un-normalising the oracle's rings by
    C(r) / (2 pi r) must give raw(r); bin 0 is raw already.  Random planes put
    power on the kx Nyquist column as well."""
    N, nz = 16, 5
    rng = np.random.default_rng(5)
    u, v, w = _fields(N, nz, rng)
    zl = [1, 2, 4]
    hats, var, ls = O.calc_spectra([(u, v, w)], zl)
    C = _ring_counts(N)
    kf = np.fft.fftfreq(N, 1.0 / N)
    rr = np.floor(np.sqrt(kf[None, :] ** 2 + kf[:, None] ** 2) + 0.5).astype(int)
    for iz, z in enumerate(zl):
        pu = 0.5 * ((u[z - 1] - u[z - 1].mean()) + (u[z] - u[z].mean()))
        pv = 0.5 * ((v[z - 1] - v[z - 1].mean()) + (v[z] - v[z].mean()))
        for f, plane in enumerate((pu, pv, w[z - 1])):
            P = np.abs(np.fft.fft2(plane) / (N * N)) ** 2
            raw = np.array([P[rr == r].sum() for r in range(N // 2 + 1)])
            got = hats[f][:, iz].copy()
            got[1:] *= C[1:] / (2 * math.pi * np.arange(1, N // 2 + 1))
            assert np.allclose(got, raw, rtol=1e-11, atol=1e-18), (f, z)


def test_spectra_pool_over_files():
    """Over two field files the ring spectra add up (nCount pools the files, and
    the factor nfiles cancels it, calc_spectra.f90:240-242), while the variances
    and integral scales are averaged over the files (:245-257)."""
    N, nz = 16, 5
    rng = np.random.default_rng(6)
    f1 = tuple(_fields(N, nz, rng))
    f2 = tuple(_fields(N, nz, rng))
    zl = [1, 3]
    h1, v1, l1 = O.calc_spectra([f1], zl)
    h2, v2, l2 = O.calc_spectra([f2], zl)
    h12, v12, l12 = O.calc_spectra([f1, f2], zl)
    for f in range(3):
        assert np.allclose(h12[f], h1[f] + h2[f], rtol=1e-12)
        assert np.allclose(v12[f], 0.5 * (v1[f] + v2[f]), rtol=1e-12)


@pytest.mark.parametrize("mode", [(3, 2), (0, 5), (7, 0), (5, 5)])
def test_spectra_single_mode(mode):
    """One Fourier mode A cos(m x + 2 n y + phi) in u (levels zl and zl+1), v and
    w: its power A^2/2 lands in ring r = nint(sqrt(m^2 + n^2)) (index radius:
    calc_spectra.f90:152 ignores xl != yl) and is scaled by
    pi ((r+1/2)^2 - (r-1/2)^2) * 3 nLevels nfiles / nCount(r), where nCount pools
    the weighted ring population over the 3 fields, the levels and the files
    (calc_spectra.f90:239-243): hat(r) = A^2/2 * 2 pi r / C(r).  Bin 0 keeps
    the raw power of w's plane mean (no normalisation).  Integral scales:
    l111 = (sum_j P(0,j) + sum_{j<N/2} P(0,j)) pi / var and l112 = sum_i P(i,0) pi / var."""
    N, nz = 16, 6
    m, nn = mode
    x, y = _xy(N, N, 2 * math.pi, math.pi)
    rng = np.random.default_rng(8)
    u, v, w = _fields(N, nz, rng)
    A = (0.7, 0.4, 0.3)
    W0 = 0.2
    zl = [2, 4]
    for z in zl:
        mode_plane = lambda amp, ph: amp * np.cos(m * x + 2 * nn * y + ph)  # noqa: E731
        u[z - 1] = mode_plane(A[0], 0.3); u[z] = mode_plane(A[0], 0.3) + 1.5   # a mean offset is removed
        v[z - 1] = mode_plane(A[1], -1.1); v[z] = mode_plane(A[1], -1.1)
        w[z - 1] = mode_plane(A[2], 0.7) + W0
    hats, var, ls = O.calc_spectra([(u, v, w)], zl)
    C = _ring_counts(N)
    r = int(math.floor(math.sqrt(m * m + nn * nn) + 0.5))
    for f in range(3):
        for iz in range(len(zl)):
            expect = np.zeros(N // 2 + 1)
            if r <= N // 2:
                expect[r] = A[f] ** 2 / 2 * 2 * math.pi * r / C[r]
            if f == 2:
                expect[0] += W0 * W0
            assert np.allclose(hats[f][:, iz], expect, rtol=1e-12, atol=1e-15), (f, iz)
            v_exp = A[f] ** 2 / 2 + (W0 * W0 if f == 2 else 0.0)
            assert math.isclose(var[f][iz], v_exp, rel_tol=1e-12)
            # integral scales from the kx = 0 column and the ky = 0 row
            P = {}
            amp2 = A[f] ** 2 / 4
            if m == 0:
                P[(0, nn)] = P.get((0, nn), 0) + amp2
                P[(0, (-nn) % N)] = P.get((0, (-nn) % N), 0) + amp2
            else:
                P[(m, nn % N)] = amp2
            if f == 2:
                P[(0, 0)] = P.get((0, 0), 0) + W0 * W0
            l1 = sum(val for (i, j), val in P.items() if i == 0) + sum(val for (i, j), val in P.items() if i == 0 and j < N // 2)
            l2 = sum(val for (i, j), val in P.items() if j == 0)
            assert math.isclose(ls[2 * f][iz], l1 * math.pi / v_exp, rel_tol=1e-12, abs_tol=1e-15), (f, "l1")
            assert math.isclose(ls[2 * f + 1][iz], l2 * math.pi / v_exp, rel_tol=1e-12, abs_tol=1e-15), (f, "l2")
