"""Pins the learned-closure oracle (oracle/sgs_mlp.py) on the CPU:
bfloat16 rounding against torch's own conversion, GELU against its
closed-form identities, the backward pass against central finite
differences, Adam against its one-step closed form, the feature stencils
against analytic fields, and the stress insertion against hand-applied rules."""
import math

import numpy as np
import pytest

import cf_oracle as O
import sgs_mlp as M


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.standard_normal(4096) * 10.0 ** rng.integers(-6, 6, 4096),
                        [0.0, 1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -1 - 2 ** -8, 65504.0, 1e-30]])
    ref = torch.tensor(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(M.bf16(x), ref)
    assert M.bf16(1 + 2 ** -8) == 1.0 and M.bf16(1 + 3 * 2 ** -8) == 1 + 2 ** -6   # ties to even


def test_gelu_identities():
    z = np.linspace(-12, 12, 20001)
    # equals the tanh form 0.5 z (1 + tanh(k (z + a z^3))) where that form does
    # not cancel (z > -3), and stays within 1e-14 of it everywhere
    tform = 0.5 * z * (1.0 + np.tanh(M.GELU_K * (z + M.GELU_A * z ** 3)))
    m = z > -3
    assert np.allclose(M.gelu(z[m]), tform[m], rtol=1e-12, atol=0)
    assert np.max(np.abs(M.gelu(z) - tform)) < 1e-14
    assert M.gelu(np.array(0.0)) == 0.0
    assert abs(M.gelu_grad(np.array(0.0)) - 0.5) < 1e-15
    assert np.allclose(M.gelu(z[z > 8]), z[z > 8]) and np.allclose(M.gelu(z[z < -8]), 0, atol=1e-12)
    # against the exact x Phi(x) within the tanh form's known accuracy
    phi = 0.5 * (1 + np.vectorize(math.erf)(z / math.sqrt(2)))
    assert np.max(np.abs(M.gelu(z) - z * phi)) < 1e-3
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - M.gelu_grad(z))) < 1e-8


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(7)
    flat = M.init_params(3) + 0.05 * rng.standard_normal(M.N_PARAMS)
    X = rng.standard_normal((37, M.N_IN))
    T = rng.standard_normal((37, M.N_OUT))
    loss, g = M.loss_and_grad(flat, X, T)
    idx = rng.choice(M.N_PARAMS, 400, replace=False)
    # every bias block and the first/last weights are in the probe set
    idx = np.unique(np.concatenate([idx, [0, 703, 704, 767, 768, 9087, 9471, 9477]]))
    h = 1e-6
    fd = np.empty(idx.size)
    for j, i in enumerate(idx):
        p = flat.copy(); p[i] += h
        lp, _ = M.loss_and_grad(p, X, T)
        p[i] -= 2 * h
        lm, _ = M.loss_and_grad(p, X, T)
        fd[j] = (lp - lm) / (2 * h)
    err = np.abs(fd - g[idx])
    assert np.max(err) < 1e-7 * max(1.0, np.max(np.abs(g))), np.max(err)


def test_bf16_mode_is_a_small_perturbation():
    rng = np.random.default_rng(8)
    flat = M.init_params(4)
    X = rng.standard_normal((256, M.N_IN))
    T = rng.standard_normal((256, M.N_OUT))
    y64, _ = M.forward(flat, X)
    y16, _ = M.forward(flat, X, bf16_mode=True)
    assert np.linalg.norm(y16 - y64) / np.linalg.norm(y64) < 2e-2
    _, g64 = M.loss_and_grad(flat, X, T)
    _, g16 = M.loss_and_grad(flat, X, T, bf16_mode=True)
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 3e-2


def test_adam_one_step_closed_form():
    rng = np.random.default_rng(9)
    p = rng.standard_normal(50)
    g = rng.standard_normal(50)
    z = np.zeros(50)
    p1, m1, v1 = M.adam(p, g, z, z, 1, lr=1e-3)
    # bias-corrected first step: m_hat = g, v_hat = g^2
    assert np.allclose(p1, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)
    assert np.allclose(m1, 0.1 * g) and np.allclose(v1, 0.001 * g * g)
    p2, _, _ = M.adam(p1, g, m1, v1, 2, lr=1e-3)
    assert np.allclose(p2, p1 - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)


def _trig_state(pr):
    """u = A(z) cos(x) + B(z) sin(2y), v = C(z) sin(x), w = D(z) cos(2y) with
    polynomial z profiles, e constant; all band-limited and periodic on
    2pi x pi."""
    n = pr.nnz
    x = np.arange(pr.nnx) * pr.dx
    y = np.arange(pr.nny) * pr.dy
    zm = (np.arange(1, n + 1) - 0.5) * pr.dz
    zw = np.arange(1, n + 1) * pr.dz
    X, Y = np.meshgrid(x, y)
    A = lambda z: 1 + 0.3 * z            # noqa: E731
    B = lambda z: 0.2 * z * z            # noqa: E731
    C = lambda z: 0.5 - 0.1 * z          # noqa: E731
    D = lambda z: 0.05 * z * (2 - z)     # noqa: E731
    u = A(zm)[:, None, None] * np.cos(X) + B(zm)[:, None, None] * np.sin(2 * Y)
    v = C(zm)[:, None, None] * np.sin(X) * np.ones_like(Y)
    w = D(zw)[:, None, None] * np.cos(2 * Y) * np.ones_like(X)
    w[-1] = 0.0
    e = np.full_like(u, 1e-3)
    return O.set_state(pr, u, v, w, e), (A, B, C, D)


def test_cell_features_closed_form():
    pr = O.Params(16, 16, 8, dt=0.01)
    st, (A, B, C, D) = _trig_state(pr)
    O.lower(st, pr); O.upper(st, pr)
    F = O.sgs_fields(st, pr)
    f = M.cell_features(F, pr)
    n = pr.nnz
    xb = np.arange(pr.nnxb) * pr.xl / pr.nnxb
    yb = np.arange(pr.nnyb) * pr.yl / pr.nnyb
    X, Y = np.meshgrid(xb, yb)
    dz = pr.dz
    for k in range(1, n + 1):
        zm = (k - 0.5) * dz
        uk = lambda kk: (0.0 if kk < 1 or kk > n else 1.0) * (  # noqa: E731
            A((kk - 0.5) * dz) * np.cos(X) + B((kk - 0.5) * dz) * np.sin(2 * Y))
        vk = lambda kk: (0.0 if kk < 1 or kk > n else 1.0) * C((kk - 0.5) * dz) * np.sin(X)  # noqa: E731
        wf = lambda kk: (0.0 if kk < 1 or kk >= n else 1.0) * D(kk * dz) * np.cos(2 * Y)  # noqa: E731
        ref = [
            -A(zm) * np.sin(X),                                   # ux
            2 * B(zm) * np.cos(2 * Y),                            # uy
            (uk(k + 1) - uk(k - 1)) / (2 * dz),                   # uz (ghosts 0)
            C(zm) * np.cos(X),                                    # vx
            0 * X,                                                # vy
            (vk(k + 1) - vk(k - 1)) / (2 * dz),                   # vz
            0 * X,                                                # wx
            None,                                                 # wy, below
            (wf(k) - wf(k - 1)) / dz,                             # wz
        ]
        ref[7] = -((D((k - 1) * dz) if k - 1 >= 1 else 0.0) + (D(k * dz) if k < n else 0.0)) * np.sin(2 * Y)
        for i in range(9):
            assert np.allclose(f[k - 1, ..., i], ref[i], atol=1e-12), (k, M.FEATURE_NAMES[i])
        assert np.allclose(f[k - 1, ..., 9], pr.dsl)
        assert np.allclose(f[k - 1, ..., 10], min(zm, pr.zl - zm))


class _Fixed:
    """Stand-in network returning a prescribed stress field."""

    def __init__(self, tau):
        self.tau = tau

    def stresses(self, feats):
        return self.tau


def test_insertion_rules_by_hand():
    pr = O.Params(16, 16, 6, dt=0.01)
    st, _ = _trig_state(pr)
    O.lower(st, pr); O.upper(st, pr)
    F = O.sgs_fields(st, pr)
    F0 = {k: v.copy() for k, v in F.items()}
    feats = M.cell_features(F, pr)
    rng = np.random.default_rng(3)
    n = pr.nnz
    tau = np.zeros(feats.shape[:-1] + (6,))
    tau[..., :3] = rng.uniform(0.5, 1.0, tau.shape[:-1] + (3,))
    tau[..., 3:] = rng.uniform(-0.1, 0.1, tau.shape[:-1] + (3,))
    ok = M.accept(tau, feats)
    assert 0.05 < ok.mean() < 0.95                   # both branches are exercised
    cl = M.LearnedClosure(_Fixed(tau))
    cl.apply(F, pr, st)
    OF = O.OF
    tr = tau[..., :3].sum(-1) / 3
    for k in range(1, n + 1):
        c = ok[k - 1]
        assert np.array_equal(F["tauuu"][k - 1], np.where(c, tau[k - 1, ..., 0] - tr[k - 1], F0["tauuu"][k - 1]))
        assert np.array_equal(F["tauuv"][k - 1], np.where(c, tau[k - 1, ..., 3], F0["tauuv"][k - 1]))
        assert np.array_equal(F["tauww"][k + OF], np.where(c, tau[k - 1, ..., 2] - tr[k - 1], F0["tauww"][k + OF]))
    for f in range(1, n):
        both = ok[f - 1] & ok[f]
        ref = np.where(both, 0.5 * (tau[f - 1, ..., 4] + tau[f, ..., 4]), F0["tauuw"][f + OF])
        assert np.array_equal(F["tauuw"][f + OF], ref)
    for f in (0, n):                                  # wall faces keep the wall model
        assert np.array_equal(F["tauuw"][f + OF], F0["tauuw"][f + OF])
        assert np.array_equal(F["tauvw"][f + OF], F0["tauvw"][f + OF])


def test_accept_criteria():
    feats = np.zeros((4, 11))
    feats[:, 1] = 1.0                                 # uy = 1: S12 = 1/2
    tau = np.array([[1, 1, 1, -0.5, 0, 0],            # realizable, -2 t12 S12 = +0.5 >= 0
                    [1, 1, 1, 0.5, 0, 0],             # realizable, backscatter
                    [1, 1, 1, -1.5, 0, 0],            # t12^2 > t11 t22
                    [1, 1, -0.1, -0.5, 0, 0]], float) # negative normal stress
    assert list(M.accept(tau, feats)) == [True, False, False, False]
    # the 3x3 determinant rule: minors fine, matrix indefinite
    t = np.array([[1, 1, 1, 0.9, 0.9, -0.9]])
    assert not M.accept(t, np.zeros((1, 11)))[0]


def test_all_fallback_equals_tke_stage():
    pr = O.Params(16, 16, 8, dt=0.01)
    u, v, w, e = O.initial_fields(pr)
    a = O.set_state(pr, u, v, w, e)
    b = O.set_state(pr, u, v, w, e)
    flat = M.init_params(1)
    flat[-6:] = [-1.0, 1, 1, 0, 0, 0]                 # t11 < 0 nearly everywhere
    L = M.unpack(flat)
    W4 = L[3][0]
    W4[:] = 0.0                                       # output = bias exactly
    flat = M.pack(L)
    net = M.Network(flat, np.zeros(11), np.ones(11), np.zeros(6), np.ones(6))
    prl = O.Params(16, 16, 8, dt=0.01, closure=M.LearnedClosure(net))
    for s in (1, 2, 3):
        O.stage(a, pr, s)
        O.stage(b, prl, s)
        assert not prl.closure.last_accept.any()
    for k in "uvwep":
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


def test_learned_stage_keeps_projection_invariants():
    pr0 = O.Params(16, 16, 8, dt=0.01)
    u, v, w, e = O.initial_fields(pr0)
    net = M.Network(M.init_params(5), np.zeros(11), np.full(11, 0.05),
                    np.array([1e-3, 1e-3, 1e-3, 0, 0, 0]), np.full(6, 1e-4))
    pr = O.Params(16, 16, 8, dt=0.01, closure=M.LearnedClosure(net))
    st = O.set_state(pr, u, v, w, e)
    O.step(st, pr)
    assert 0 < pr.closure.last_accept.mean() < 1
    div = O.divergence(st, pr)
    scale = np.abs(O.xderiv(O.interior(st, pr, "u"), pr.xl)).max()
    assert np.abs(div).max() < 1e-12 * scale
    wm = O.interior(st, pr, "w")[:-1].mean(axis=(1, 2))
    assert np.abs(wm).max() < 1e-15


# ----------------------------------------------------------------------------
# small-grid mode
# ----------------------------------------------------------------------------
def test_small_grid_all_fallback_equals_tke_on_the_small_grid():
    """With every cell rejected, the small-grid mode is the TKE closure with
    its stresses evaluated on the small grid.  Against the 3/2-grid TKE stage
    that is a different discretisation: close, not equal."""
    pr0 = O.Params(16, 16, 8, dt=0.01)
    u, v, w, e = O.initial_fields(pr0)
    flat = M.init_params(1)
    L = M.unpack(flat)
    L[3][0][:] = 0.0
    L[3][1][:] = [-1.0, 1, 1, 0, 0, 0]
    net = M.Network(M.pack(L), np.zeros(11), np.ones(11), np.zeros(6), np.ones(6))
    cl = M.LearnedClosureSmall(net)
    pr = O.Params(16, 16, 8, dt=0.01, closure=cl)
    a = O.set_state(pr0, u, v, w, e)
    b = O.set_state(pr, u, v, w, e)
    O.stage(a, pr0, 1)
    O.stage(b, pr, 1)
    assert not cl.last_accept.any()
    for k in "uvw":
        r = np.linalg.norm(O.interior(b, pr, k) - O.interior(a, pr0, k)) / np.linalg.norm(O.interior(a, pr0, k))
        assert 0 < r < 1e-3, (k, r)


def test_small_grid_stresses_by_hand_at_one_point():
    """The TKE fallback stresses of the small-grid mode at one point, from the
    formulas of compsgs1.f90:41-50 written out directly."""
    pr = O.Params(16, 16, 8, dt=0.01)
    u, v, w, e = O.initial_fields(pr)
    st = O.set_state(pr, u, v, w, e)
    O.lower(st, pr); O.upper(st, pr)
    flat = M.init_params(1)
    L = M.unpack(flat)
    L[3][0][:] = 0.0
    L[3][1][:] = [-1.0, 1, 1, 0, 0, 0]
    net = M.Network(M.pack(L), np.zeros(11), np.ones(11), np.zeros(6), np.ones(6))
    (c11, c22, c12, c33), (f13, f23), ok = M.small_grid_stresses(st, pr, net)
    k, j, i = 3, 5, 7                                   # level 3, row 5, column 7
    lv = lambda a, L: O.lev(a, L)                       # noqa: E731
    K = lambda f: min(0.1 * pr.alk * math.sqrt(lv(st.e, f)[j, i]), pr.xkmax)   # noqa: E731
    ux = O.xderiv(lv(st.u, k), pr.xl)[j, i]
    wz = (lv(st.w, k) - lv(st.w, k - 1))[j, i] / pr.dz
    assert abs(c11[k - 1, j, i] - (-(K(k - 1) + K(k)) * ux)) < 1e-14
    assert abs(c33[k - 1, j, i] - (-(K(k - 1) + K(k)) * wz)) < 1e-14
    uzf = (lv(st.u, k + 1) - lv(st.u, k))[j, i] / pr.dz
    wxf = O.xderiv(lv(st.w, k), pr.xl)[j, i]
    assert abs(f13[k - 1, j, i] - (-K(k) * (uzf + wxf))) < 1e-14


def test_small_grid_stage_keeps_projection_invariants():
    pr0 = O.Params(16, 16, 8, dt=0.01)
    u, v, w, e = O.initial_fields(pr0)
    net = M.Network(M.init_params(5), np.zeros(11), np.full(11, 0.05),
                    np.array([1e-3, 1e-3, 1e-3, 0, 0, 0]), np.full(6, 1e-4))
    cl = M.LearnedClosureSmall(net)
    pr = O.Params(16, 16, 8, dt=0.01, closure=cl)
    st = O.set_state(pr, u, v, w, e)
    O.step(st, pr, mnout=True)
    assert 0 < cl.last_accept.mean() < 1
    div = O.divergence(st, pr)
    scale = np.abs(O.xderiv(O.interior(st, pr, "u"), pr.xl)).max()
    assert np.abs(div).max() < 1e-12 * scale
    assert np.abs(O.interior(st, pr, "w")[:-1].mean(axis=(1, 2))).max() < 1e-15
