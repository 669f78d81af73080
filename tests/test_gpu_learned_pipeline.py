"""The learned-closure training pipeline end to end on the GPU, at a size that
runs in seconds: a fine FP32 run, the device box filter, the coarse samples,
several candidate networks trained on the same samples by the library's
kernels, their a-posteriori validation on both closure grids, the selection,
and the saved network file, which must load and reproduce the kept
candidate's held-out error.  At this size (20 spin-up steps) the networks
are accepted at no cell of the validation runs, so the selection takes its
fallback (the lowest score of all); tests/test_learned_pipeline.py covers
the rule's other branches."""
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "paes-channelflow_amd"))


def test_pipeline_trains_validates_and_keeps_a_candidate(tmp_path):
    import channelflow as C
    from channelflow import learned as L
    out = tmp_path / "net"
    L.main(["--nx", "128", "--nz", "64", "--spinup", "20", "--snaps", "3", "--every", "5",
            "--filters", "4,4,2", "--batch", "1024", "--schedule", "300:0.001",
            "--candidates", "3", "--val-grid", "128,32", "--val-steps", "6", "--fixture", "512",
            "--out-dir", str(out)])
    meta = json.load(open(out / "sgs_mlp.json"))
    cands = meta["candidates"]
    assert len(cands) == 3 and meta["min_accept"] == 0.3
    # 128^2 x 64 filtered by 4 x 4 x 2: 32^2 x 32 samples per snapshot, two kept
    assert meta["train_samples"] == 2 * 32 * 32 * 32 and meta["held_out_samples"] == 32 * 32 * 32
    for cd in cands:
        v = cd["validation"]
        assert set(v) >= {"small", "dealiased", "score", "accepted"}
        for g in ("small", "dealiased"):
            assert 0.0 <= v[g]["accepted"] <= 1.0 and np.isfinite(v[g]["uxym"])
        assert v["score"] == max(v["small"]["score"], v["dealiased"]["score"])
        assert v["accepted"] == min(v["small"]["accepted"], v["dealiased"]["accepted"])
        assert np.isfinite(cd["held_out"]["mse_dev"]["mlp"]) and cd["final_loss"] > 0
    best = meta["selected"]
    assert best == L.select_candidate([cd["validation"] for cd in cands], meta["min_accept"])
    assert meta["held_out"] == cands[best]["held_out"]
    # the saved file is the kept candidate: its predictions on the held-out
    # fixture give the recorded a-priori error on those samples' subset
    fx = np.load(out / "sgs_heldout.npz")
    net = C.SgsNet(0, str(out / "sgs_mlp.cfmlp"))
    try:
        pred = net.forward(fx["feats"])
    finally:
        net.close()
    assert pred.shape == (fx["feats"].shape[0], 6) and np.all(np.isfinite(pred))
    mse = L._dev_err(pred, fx["tau"])
    full = cands[best]["held_out"]["mse_dev"]["mlp"]
    assert 0.2 * full < mse < 5.0 * full, (mse, full)
