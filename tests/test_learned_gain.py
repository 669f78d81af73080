"""What the trained network gains over the closure without it, a priori.

The network is the one shipped in paes-channelflow_amd/weights/sgs_mlp.cfmlp,
trained on the device by the library's kernels.  It is evaluated on held-out
samples: 16384 cells of the last snapshot of the 512^2 x 256 FP32 run
(Re_tau 395), box-filtered 4x4x2 and 4x4x4, which training never saw
(tests/golden/sgs_heldout_512x256.npz, written by channelflow.learned).  The
targets are the complete subfilter stresses: the resolved part plus the fine
run's own SGS stress.  Baselines:
- the TKE closure -2 K S with K = 0.1 dsl sqrt(e) (complen.f90:20-21);
- the same with its best least-squares coefficient;
- no model.
The CPU test uses the FP64 oracle network with BF16 emulation; the GPU test
runs the device inference kernel on the same samples.
"""
import os

import numpy as np
import pytest

import sgs_mlp as M
from sgs_weights import read_cfmlp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "paes-channelflow_amd", "weights", "sgs_mlp.cfmlp")
HELD = os.path.join(ROOT, "tests", "golden", "sgs_heldout_512x256.npz")
W6 = np.array([1, 1, 1, 2, 2, 2], float)


def _dev(t):
    d = np.array(t, float)
    d[:, :3] -= d[:, :3].sum(1, keepdims=True) / 3
    return d


def _mse(pred, true):
    x = _dev(pred) - _dev(true)
    return float(np.mean((x * x) @ W6))


def _diss(t, f):
    S = np.stack([f[:, 0], f[:, 4], f[:, 8], 0.5 * (f[:, 1] + f[:, 3]), 0.5 * (f[:, 2] + f[:, 6]),
                  0.5 * (f[:, 5] + f[:, 7])], -1)
    return -(_dev(t) * S) @ W6


def _load():
    d = np.load(HELD)
    return d["feats"].astype(float), d["tau"].astype(float), d["e_sgs"].astype(float)


def _baselines(f, t, e):
    tke = M.tke_stress_apriori(f, e, f[:, 9])
    dt, dk = _dev(t), _dev(tke)
    c = ((dt * dk) @ W6).sum() / ((dk * dk) @ W6).sum()
    return tke, {"tke": _mse(tke, t), "tke_best": _mse(c * tke, t), "zero": _mse(0 * t, t)}


def _check_gain(pred, f, t, e):
    tke, base = _baselines(f, t, e)
    m = _mse(pred, t)
    print("a-priori deviatoric MSE: mlp", m, base, flush=True)
    # the network halves the error of the best eddy-viscosity fit and of no model
    assert m < 0.6 * min(base.values())
    # and gets the mean SGS dissipation within 25 %, where the TKE closure is 2x off
    true_eps = _diss(t, f).mean()
    assert abs(_diss(pred, f).mean() - true_eps) < 0.25 * abs(true_eps)
    assert abs(_diss(tke, f).mean() - true_eps) > 0.5 * abs(true_eps)


def test_trained_network_beats_the_closure_without_it():
    flat, imu, isd, omu, osd = read_cfmlp(WEIGHTS)
    f, t, e = _load()
    assert set(np.round(np.unique(f[:, 9]), 6)) == {0.026602, 0.033516}   # both filter widths
    pred = M.Network(flat, imu, isd, omu, osd, bf16_mode=True).stresses(f)
    _check_gain(pred, f, t, e)


@pytest.mark.gpu
def test_device_network_reproduces_the_gain():
    import channelflow as C
    flat, imu, isd, omu, osd = read_cfmlp(WEIGHTS)
    f, t, e = _load()
    net = C.SgsNet(0, WEIGHTS)
    pred = net.forward(f.astype(np.float32)).astype(float)
    net.close()
    ref = M.Network(flat, imu, isd, omu, osd, bf16_mode=True).stresses(f)
    err = np.linalg.norm(pred - ref) / np.linalg.norm(ref - omu)
    print("device vs oracle network", err, flush=True)
    assert err < 5e-3
    _check_gain(pred, f, t, e)
