"""FP64 oracle of the learned SGS closure (no counterpart in the reference).

TEST INFRASTRUCTURE ONLY.  This module is a checker: it may be imported by
``tests/`` and ``__graft_entry__.smoke()`` and by nothing else.  The product
path (the HIP trainer and the closure inside the row-tendency kernel) never
calls it.

What it restates:

* the network: a pointwise MLP 11 -> 64 -> 64 -> 64 -> 6, GELU (tanh form,
  0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)))) after each hidden layer,
  linear output, on standardised inputs and outputs;
* its gradient (backward pass) and the Adam update;
* the device numerics, on request (``bf16=True``).  This emulates the BF16
  MFMA path of the HIP kernels:
  - the normalised inputs, the weights, the post-GELU activations and the
    back-propagated output/pre-activation gradients are rounded to bfloat16
    (round to nearest even);
  - every product accumulates at higher precision;
  - the bias gradients are sums of the rounded gradients;
* the features the closure sees in the LES, at cell centres of the 3/2 grid:
  - ux, uy, vx, vy: spectral derivatives at the level;
  - uz, vz = (u(k+1) - u(k-1)) / (2 dz), with the ghost levels of lower/upper;
  - wx, wy = the mean of the two faces' values;
  - wz = (w(k) - w(k-1)) / dz;
  - the filter width dsl (les.F90:118);
  - the wall distance min(zm, zl - zm);
* the insertion at the stress computation of compsgs1.f90:1-53.  The
  stresses sit at the same staggered locations, with the TKE closure as
  fallback:
  - a cell uses the network's deviatoric stress when the predicted stress is
    realizable (non-negative principal minors) and dissipative
    (-tau_ij S_ij >= 0); otherwise it uses the TKE stress;
  - a face (tau13, tau23) uses the mean of its two cells' network values when
    both cells passed, and the TKE face stress otherwise;
  - the wall faces keep the wall model (lowerbc.f90:129-130,
    upperbc.f90:127-128).

Pinning: the backward pass against central finite differences of the FP64
forward; Adam against its closed form after one step (m/(sqrt(v)+eps) =
sign(g) for bias-corrected moments); GELU against its closed form and
math.erf-free identities; the feature stencils against closed-form
trigonometric fields; the insertion against the TKE stage when every cell
falls back and against a constant-stress network
(tests/test_oracle_sgs_mlp.py).
"""
from __future__ import annotations

import math

import numpy as np

import cf_oracle as O

N_IN, N_HID, N_OUT = 11, 64, 6
LAYERS = ((N_HID, N_IN), (N_HID, N_HID), (N_HID, N_HID), (N_OUT, N_HID))
# flat parameter order: W1 (64x11, row-major [out][in]), b1, W2, b2, W3, b3, W4 (6x64), b4
N_PARAMS = sum(o * i + o for o, i in LAYERS)        # 9478
FEATURE_NAMES = ("ux", "uy", "uz", "vx", "vy", "vz", "wx", "wy", "wz", "delta", "wall")
STRESS_NAMES = ("t11", "t22", "t33", "t12", "t13", "t23")
GELU_K = math.sqrt(2.0 / math.pi)
GELU_A = 0.044715


def bf16(a):
    """Round to bfloat16 (nearest even) through float32, as the device does."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def _sigmoid(x):
    e = np.exp(-np.abs(x))
    return np.where(x >= 0, 1.0 / (1.0 + e), e / (1.0 + e))


def gelu(z):
    """0.5 z (1 + tanh(k (z + a z^3))) written as z * sigmoid(2 k (z + a z^3)),
    which does not cancel for negative z."""
    return z * _sigmoid(2.0 * GELU_K * (z + GELU_A * z ** 3))


def gelu_grad(z):
    s = _sigmoid(2.0 * GELU_K * (z + GELU_A * z ** 3))
    return s + z * s * (1.0 - s) * 2.0 * GELU_K * (1.0 + 3.0 * GELU_A * z * z)


def unpack(flat):
    flat = np.asarray(flat, dtype=np.float64)
    assert flat.size == N_PARAMS
    out, o = [], 0
    for (m, n) in LAYERS:
        W = flat[o: o + m * n].reshape(m, n); o += m * n
        b = flat[o: o + m]; o += m
        out.append((W, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b]) for W, b in layers])


def init_params(seed=0):
    """He-uniform-like initial weights, zero biases (the trainer's default)."""
    rng = np.random.default_rng(seed)
    layers = []
    for (m, n) in LAYERS:
        lim = math.sqrt(6.0 / n) * (0.5 if m == N_OUT else 1.0)
        layers.append((rng.uniform(-lim, lim, (m, n)), np.zeros(m)))
    return pack(layers)


def forward(flat, X, bf16_mode=False):
    """Y = MLP(X) for normalised inputs X (n, 11); returns (Y, cache)."""
    L = unpack(flat)
    r = bf16 if bf16_mode else (lambda a: np.asarray(a, dtype=np.float64))
    H = r(X)
    cache = {"H0": H, "Z": []}
    for li, (W, b) in enumerate(L):
        Z = H @ r(W).T + b
        if li == len(L) - 1:
            return Z, cache
        cache["Z"].append(Z)
        H = r(gelu(Z))
        cache[f"H{li + 1}"] = H


def backward(flat, cache, dY, bf16_mode=False):
    """Gradient of sum(dY * Y) with respect to the flat parameters."""
    L = unpack(flat)
    r = bf16 if bf16_mode else (lambda a: np.asarray(a, dtype=np.float64))
    grads = [None] * len(L)
    dZ = r(dY)
    for li in range(len(L) - 1, -1, -1):
        W, _ = L[li]
        Hin = cache[f"H{li}"]
        grads[li] = (dZ.T @ Hin, dZ.sum(axis=0))
        if li == 0:
            break
        dH = dZ @ r(W)
        dZ = r(dH * gelu_grad(cache["Z"][li - 1]))
    return pack(grads)


def loss_and_grad(flat, X, T, bf16_mode=False):
    """Mean squared error over samples and the 6 outputs, and its gradient."""
    Y, cache = forward(flat, X, bf16_mode)
    R = Y - T
    n = X.shape[0]
    loss = float(np.sum(R * R) / (n * N_OUT))
    g = backward(flat, cache, 2.0 * R / (n * N_OUT), bf16_mode)
    return loss, g


def adam(flat, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t counts from 1); returns new (flat, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return flat - lr * mh / (np.sqrt(vh) + eps), m, v


# ----------------------------------------------------------------------------
# Normalisation and the network as the closure sees it
# ----------------------------------------------------------------------------
class Network:
    """Parameters plus the input/output standardisation:
    x = (f - in_mu) / in_sd, tau = y * out_sd + out_mu."""

    def __init__(self, flat, in_mu, in_sd, out_mu, out_sd, bf16_mode=True):
        self.flat = np.asarray(flat, dtype=np.float64)
        self.in_mu = np.asarray(in_mu, dtype=np.float64)
        self.in_sd = np.asarray(in_sd, dtype=np.float64)
        self.out_mu = np.asarray(out_mu, dtype=np.float64)
        self.out_sd = np.asarray(out_sd, dtype=np.float64)
        self.bf16_mode = bf16_mode

    def stresses(self, feats):
        """feats (..., 11) raw -> tau (..., 6) in STRESS_NAMES order."""
        shp = feats.shape[:-1]
        F = feats.reshape(-1, N_IN)
        X = (F - self.in_mu) / self.in_sd
        if self.bf16_mode:
            X = X.astype(np.float32).astype(np.float64)   # normalised in FP32 on the device
        Y, _ = forward(self.flat, X, self.bf16_mode)
        return (Y * self.out_sd + self.out_mu).reshape(shp + (N_OUT,))


def cell_features(F, pr: O.Params):
    """The 11 closure inputs at cell centres of levels 1..nnz on the 3/2 grid,
    from the comp1 intermediates of cf_oracle.sgs_fields.  Shape (nnz, nyb, nxb, 11)."""
    n = pr.nnz
    OF = O.OF
    sl = slice(1 + OF, n + 1 + OF)
    km = slice(0 + OF, n + OF)
    ub, vb = F["ub"], F["vb"]
    uz = (ub[2 + OF: n + 2 + OF] - ub[km]) * pr.todzm
    vz = (vb[2 + OF: n + 2 + OF] - vb[km]) * pr.todzm
    wx = 0.5 * (F["wx"][km] + F["wx"][sl])
    wy = 0.5 * (F["wy"][km] + F["wy"][sl])
    zm = (np.arange(1, n + 1) - 0.5) * pr.dz
    wall = np.minimum(zm, pr.zl - zm)[:, None, None] * np.ones_like(uz)
    delta = np.full_like(uz, pr.dsl)
    return np.stack([F["ux"][sl], F["uy"][sl], uz, F["vx"][sl], F["vy"][sl], vz, wx, wy,
                     F["wz"][sl], delta, wall], axis=-1)


def accept(tau, feats):
    """Cells where the predicted stress is realizable (every principal minor
    non-negative) and dissipative (-tau_ij S_ij >= 0)."""
    t11, t22, t33, t12, t13, t23 = (tau[..., i] for i in range(6))
    ux, uy, uz, vx, vy, vz, wx, wy, wz = (feats[..., i] for i in range(9))
    det = (t11 * (t22 * t33 - t23 * t23) - t12 * (t12 * t33 - t23 * t13)
           + t13 * (t12 * t23 - t22 * t13))
    real = ((t11 >= 0) & (t22 >= 0) & (t33 >= 0) & (t12 * t12 <= t11 * t22)
            & (t13 * t13 <= t11 * t33) & (t23 * t23 <= t22 * t33) & (det >= 0))
    eps = -(t11 * ux + t22 * vy + t33 * wz + t12 * (uy + vx) + t13 * (uz + wx) + t23 * (vz + wy))
    return real & (eps >= 0)


class LearnedClosure:
    """The closure hook of cf_oracle.comp1: replaces the stresses of
    compsgs1 with the network's, TKE where a cell fails ``accept``."""

    def __init__(self, net: Network):
        self.net = net
        self.last_accept = None

    def apply(self, F, pr: O.Params, st):
        n = pr.nnz
        OF = O.OF
        sl = slice(1 + OF, n + 1 + OF)
        feats = cell_features(F, pr)
        tau = self.net.stresses(feats)
        ok = accept(tau, feats)
        self.last_accept = ok
        tr = (tau[..., 0] + tau[..., 1] + tau[..., 2]) / 3.0
        F["tauuu"] = np.where(ok, tau[..., 0] - tr, F["tauuu"])
        F["tauvv"] = np.where(ok, tau[..., 1] - tr, F["tauvv"])
        F["tauuv"] = np.where(ok, tau[..., 3], F["tauuv"])
        tww = F["tauww"].copy()
        tww[sl] = np.where(ok, tau[..., 2] - tr, F["tauww"][sl])
        F["tauww"] = tww
        # faces 1..nnz-1 between cells f and f+1 (array index f-1 and f)
        okf = ok[:-1] & ok[1:]
        fs = slice(1 + OF, n + OF)
        tuw = F["tauuw"].copy()
        tvw = F["tauvw"].copy()
        tuw[fs] = np.where(okf, 0.5 * (tau[:-1, ..., 4] + tau[1:, ..., 4]), F["tauuw"][fs])
        tvw[fs] = np.where(okf, 0.5 * (tau[:-1, ..., 5] + tau[1:, ..., 5]), F["tauvw"][fs])
        F["tauuw"] = tuw
        F["tauvw"] = tvw


def tke_stress_apriori(feats, e_sgs, alk, coef=0.1):
    """Deviatoric TKE-closure stress -2 K S_ij with K = coef*alk*sqrt(e)
    (complen.f90:20-21, compsgs1.f90:41-50), in STRESS_NAMES order."""
    ux, uy, uz, vx, vy, vz, wx, wy, wz = (feats[..., i] for i in range(9))
    K = coef * alk * np.sqrt(np.maximum(e_sgs, 0.0))
    return np.stack([-2 * K * ux, -2 * K * vy, -2 * K * wz, -K * (uy + vx), -K * (uz + wx),
                     -K * (vz + wy)], axis=-1)


def deviatoric(tau):
    tr = (tau[..., 0] + tau[..., 1] + tau[..., 2]) / 3.0
    d = tau.copy()
    for i in range(3):
        d[..., i] -= tr
    return d


# ----------------------------------------------------------------------------
# The closure evaluated on the nnx x nny grid ("small grid" mode)
# ----------------------------------------------------------------------------
def small_grid_stresses(st, pr: O.Params, net):
    """Network and TKE stresses at the cell centres of the small grid, from the
    stage-start state (after lower/upper).  Returns (cell, faces, ok):
    cell = (t11, t22, t12, t33) at levels 1..nnz; faces = (t13, t23) at
    faces 1..nnz-1; ok = accepted cells."""
    n = pr.nnz
    OF = O.OF
    xl, yl = pr.xl, pr.yl
    u = st.u[0 + OF: n + 2 + OF]                     # levels 0..nnz+1
    v = st.v[0 + OF: n + 2 + OF]
    w = st.w[0 + OF: n + 1 + OF]                     # faces 0..nnz
    e = st.e[0 + OF: n + 1 + OF].copy()              # faces 0..nnz
    e[0] = e[1]                                      # lower (lowerbc.f90:15)
    e[n] = e[n - 1]                                  # upper (upperbc.f90:20-24)
    uk, vk = u[1:n + 1], v[1:n + 1]
    ux, uy = O.xderiv(uk, xl), O.yderiv(uk, yl)
    vx, vy = O.xderiv(vk, xl), O.yderiv(vk, yl)
    uz = (u[2:n + 2] - u[0:n]) * pr.todzm
    vz = (v[2:n + 2] - v[0:n]) * pr.todzm
    wxf, wyf = O.xderiv(w, xl), O.yderiv(w, yl)
    wx, wy = 0.5 * (wxf[:-1] + wxf[1:]), 0.5 * (wyf[:-1] + wyf[1:])
    wz = (w[1:] - w[:-1]) * pr.dzm
    zm = (np.arange(1, n + 1) - 0.5) * pr.dz
    wall = np.minimum(zm, pr.zl - zm)[:, None, None] * np.ones_like(uz)
    feats = np.stack([ux, uy, uz, vx, vy, vz, wx, wy, wz, np.full_like(uz, pr.dsl), wall], axis=-1)
    K = np.minimum(0.1 * pr.alk * np.sqrt(e), pr.xkmax)          # complen.f90:20-21, faces 0..nnz
    xs = K[:-1] + K[1:]
    tke = (-xs * ux, -xs * vy, -xs * 0.5 * (uy + vx), -xs * wz)   # compsgs1.f90:41-48
    uzf = (u[2:n + 1] - u[1:n]) * pr.dzm                          # faces 1..nnz-1
    vzf = (v[2:n + 1] - v[1:n]) * pr.dzm
    tke13 = -K[1:n] * (uzf + wxf[1:n])                            # compsgs1.f90:49-50
    tke23 = -K[1:n] * (vzf + wyf[1:n])
    tau = net.stresses(feats)
    ok = accept(tau, feats)
    tr = (tau[..., 0] + tau[..., 1] + tau[..., 2]) / 3.0
    cell = (np.where(ok, tau[..., 0] - tr, tke[0]), np.where(ok, tau[..., 1] - tr, tke[1]),
            np.where(ok, tau[..., 3], tke[2]), np.where(ok, tau[..., 2] - tr, tke[3]))
    okf = ok[:-1] & ok[1:]
    faces = (np.where(okf, 0.5 * (tau[:-1, ..., 4] + tau[1:, ..., 4]), tke13),
             np.where(okf, 0.5 * (tau[:-1, ..., 5] + tau[1:, ..., 5]), tke23))
    return cell, faces, ok


class LearnedClosureSmall:
    """The closure hook of cf_oracle.comp1 in small-grid mode.

    It does three things:
    - The interior SGS momentum stresses of the 3/2-grid sweep are switched
      off; the wall-model faces 0 and nnz stay there.
    - The stresses of ``small_grid_stresses`` are differentiated on the small
      grid: spectral d/dx, d/dy; dz differences with zero interior flux at
      the walls.
    - The result is band-limited (Nyquist modes dropped) and added to the
      compressed tendencies; the plane mean of the w tendency is removed, as
      comp1.f90:172 does for rw.
    """

    def __init__(self, net: Network):
        self.net = net
        self.last_accept = None
        self._r = None
        self._means = None

    def apply(self, F, pr: O.Params, st):
        n = pr.nnz
        OF = O.OF
        xl, yl = pr.xl, pr.yl
        (c11, c22, c12, c33), (f13, f23), ok = small_grid_stresses(st, pr, self.net)
        self.last_accept = ok
        shp = c11.shape[1:]
        zero = np.zeros((1,) + shp)
        F13 = np.concatenate([zero, f13, zero])                  # faces 0..nnz
        F23 = np.concatenate([zero, f23, zero])
        ru = -O.xderiv(c11, xl) - O.yderiv(c12, yl) - (F13[1:] - F13[:-1]) * pr.dzm
        rv = -O.xderiv(c12, xl) - O.yderiv(c22, yl) - (F23[1:] - F23[:-1]) * pr.dzm
        rw = -O.xderiv(f13, xl) - O.yderiv(f23, yl) - (c33[1:] - c33[:-1]) * pr.dzm
        rw = O.deal(rw)
        rw -= rw.mean(axis=(-2, -1), keepdims=True)
        self._r = (O.deal(ru), O.deal(rv), rw)
        mean = lambda a: a.mean(axis=(-2, -1))                   # noqa: E731
        self._means = dict(uusb=mean(c11), vvsb=mean(c22), uvsb=mean(c12), wwsb=mean(c33),
                           uwsb=np.concatenate([mean(f13), [0.0]]), vwsb=np.concatenate([mean(f23), [0.0]]))
        # the 3/2-grid sweep keeps only the wall stresses
        for k in ("tauuu", "tauvv", "tauuv"):
            F[k] = np.zeros_like(F[k])
        F["tauww"] = np.zeros_like(F["tauww"])
        for k in ("tauuw", "tauvw"):
            t = F[k].copy()
            t[1 + OF: n + OF] = 0.0
            F[k] = t

    def post(self, st, pr: O.Params, means: bool):
        n = pr.nnz
        OF = O.OF
        ru, rv, rw = self._r
        st.ru[1 + OF: n + 1 + OF] += ru
        st.rv[1 + OF: n + 1 + OF] += rv
        st.rw[1 + OF: n + OF] += rw
        if means:                                  # the SGS means of compsgs3.f90:71-78
            for k, v in self._means.items():
                st.sgs[k] = st.sgs[k] + v
