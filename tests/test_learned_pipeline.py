"""The training-data side of the learned closure on the CPU: the coarse-grid
features of channelflow.learned use the same stencils as the closure in the
LES.  That is checked against the oracle's 3/2-grid cell features at the grid
points the two grids share, for a band-limited state.  Also checked: the
subfilter-stress targets, the a-priori metrics on constructed cases, and the
rule that picks a candidate network by its a-posteriori validation score."""
import math
import sys
import os

import numpy as np
import pytest

import cf_oracle as O
import sgs_mlp as M

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "paes-channelflow_amd"))


def _learned():
    import importlib
    return importlib.import_module("channelflow.learned")


def test_coarse_features_equal_les_features_at_shared_points():
    L = _learned()
    pr = O.Params(16, 16, 8, dt=0.01)
    u, v, w, e = O.initial_fields(pr)
    st = O.set_state(pr, u, v, w, e)       # deal: band-limited
    O.lower(st, pr); O.upper(st, pr)
    F = O.sgs_fields(st, pr)
    ref = M.cell_features(F, pr)           # (nz, nyb, nxb, 11) on the 3/2 grid
    n = pr.nnz
    fd = {"u": O.interior(st, pr, "u"), "v": O.interior(st, pr, "v"), "wf": O.interior(st, pr, "w"),
          "wc": np.zeros_like(u), "e": np.zeros_like(u)}
    for k in ("uu", "vv", "ww", "uv", "uw", "vw"):
        fd[k] = np.zeros_like(u)
    feats, _, _ = L.coarse_samples(fd, pr.xl, pr.yl, pr.zl)
    feats = feats.reshape(n, pr.nny, pr.nnx, 11).astype(np.float64)
    # small-grid point (j, i) = 3/2-grid point (3j/2, 3i/2) for even j, i
    a = feats[:, ::2, ::2, :]
    b = ref[:, ::3, ::3, :]
    scale = np.abs(b[..., :9]).max()
    assert np.max(np.abs(a[..., :9] - b[..., :9])) < 1e-6 * scale   # features are FP32
    assert np.allclose(a[..., 9], pr.dsl, rtol=1e-6) and np.allclose(a[..., 10], b[..., 10], rtol=1e-6)


def test_targets_are_subfilter_stresses():
    L = _learned()
    rng = np.random.default_rng(0)
    shp = (4, 8, 8)
    fd = {k: rng.standard_normal(shp) for k in ("u", "v", "wc", "wf", "e")}
    fd["wf"][-1] = 0.0
    prods = {"uu": ("u", "u"), "vv": ("v", "v"), "ww": ("wc", "wc"), "uv": ("u", "v"),
             "uw": ("u", "wc"), "vw": ("v", "wc")}
    for k, (a, b) in prods.items():
        fd[k] = fd[a] * fd[b] + 0.5 + 0.1 * rng.random(shp)
    _, tau, es = L.coarse_samples(fd, 2 * math.pi, math.pi, 2.0)
    for i, k in enumerate(("uu", "vv", "ww", "uv", "uw", "vw")):
        a, b = prods[k]
        assert np.allclose(tau[:, i], (fd[k] - fd[a] * fd[b]).reshape(-1), atol=1e-6)
    assert np.allclose(es, 0.5 * tau[:, :3].sum(1) + fd["e"].reshape(-1), atol=1e-6)


def test_apriori_metrics_on_constructed_cases():
    L = _learned()
    rng = np.random.default_rng(1)
    n = 4000
    feats = rng.standard_normal((n, 11)).astype(np.float32)
    feats[:, 9] = 0.05
    e = rng.uniform(0.1, 1.0, n).astype(np.float32)
    tke = L.tke_stress(feats, e)
    tau = (2.0 * tke).astype(np.float32)          # truth = twice the TKE closure
    ev = L.evaluate(tau.copy(), feats, tau, e)    # a perfect model
    assert ev["mse_dev"]["mlp"] < 1e-12 * ev["mse_dev"]["zero"]
    assert abs(ev["tke_best_coef_scale"] - 2.0) < 1e-5
    assert ev["mse_dev"]["tke_best_coef"] < 1e-10 * ev["mse_dev"]["zero"]
    assert abs(ev["dissipation_corr"]["tke"] - 1.0) < 1e-9
    # the TKE closure at coefficient 0.1 is the oracle's tke_stress_apriori
    ref = M.tke_stress_apriori(feats.astype(float), e.astype(float), 0.05)
    assert np.allclose(tke, ref, rtol=1e-6, atol=1e-9)


def test_candidate_selection_rule():
    """The lowest validation score among candidates accepted at min_accept
    or more; a candidate that falls back almost everywhere cannot win on a
    score it gets by matching the TKE closure; with none above the floor the
    lowest score of all; a diverged candidate (score inf) never wins."""
    L = _learned()
    v = lambda acc, score: {"accepted": acc, "score": score}   # noqa: E731
    vals = [v(0.45, 2.7), v(0.09, 0.2), v(0.42, 0.25), v(0.50, float("inf")), v(0.41, 0.45)]
    assert L.select_candidate(vals, 0.3) == 2
    assert L.select_candidate(vals, 0.0) == 1
    assert L.select_candidate([v(0.1, 3.0), v(0.2, 1.5), v(0.05, 0.5)], 0.3) == 2
    assert L.select_candidate([v(0.4, float("inf")), v(0.1, 9.0)], 0.3) == 1


def test_pipeline_refuses_zero_candidates():
    L = _learned()
    with pytest.raises(SystemExit):
        L.main(["--candidates", "0", "--out-dir", "unused"])
