"""Training pipeline of the learned SGS closure (no counterpart in the reference).

The samples come from this code's own finer-grid run:

1. A fine FP32 run of the channel, from the seeded log-law initial state.
2. Every few steps after spin-up, the device box filter (cf_filter_snapshot)
   averages the state over fx x fy x fz fine cells.
3. The coarse fields give the 11 closure inputs, with the stencils the LES
   itself uses at a cell centre (the same as k_rows_mlp):
   - spectral ux uy vx vy of the band-limited (Nyquist-free) coarse field;
   - uz, vz = (u(K+1) - u(K-1)) / (2 dz) with zero ghost levels;
   - wx, wy = the mean of the two coarse faces;
   - wz = (w(K) - w(K-1)) / dz;
   - the coarse filter width (dx dy dz)^(1/3);
   - the wall distance.
4. The targets are the complete subfilter stresses: the resolved part
   <u_i u_j> - <u_i><u_j> of the cell-centred fine velocities, plus the box
   mean of the fine run's own TKE-closure stresses.
5. The network is trained on the device by the library's BF16 MFMA forward,
   backward and Adam kernels (SgsNet.train).  Several candidates are trained
   on the same samples from different initial weights and batch orders.
6. Each candidate is run a posteriori: the LES of the validation case (the
   learned-closure grid, FP32, from a seeded state whose seed differs from
   the one the tests use), once with the network on the nnx x nny grid and
   once on the 3/2 grid, against the FP64 TKE-closure run of the same state.
   Candidates with the same a-priori error differ widely here, so the one
   with the smallest validation score is kept: over both closure grids, the
   larger of its mean-profile error over 1 % and its resolved-TKE deviation
   over 5 %.  Only candidates whose stresses are
   accepted at a fraction --min-accept of the cells or more compete (one that
   falls back to the TKE closure nearly everywhere would match that closure
   trivially); if none is, all do.

``evaluate`` measures a trained network a priori on held-out samples.  It
compares the deviatoric stress error against three baselines:
- the TKE closure's -2 K S_ij with K = 0.1 dsl sqrt(e) (complen.f90:20-21);
- that closure with its best least-squares coefficient;
- no model.
Here e is the resolved subfilter energy plus the filtered fine-grid e.

Run ``python -m channelflow.learned --help`` (from paes-channelflow_amd/).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import (CF_CLOSURE_GRID_DEALIASED, CF_CLOSURE_GRID_SMALL, CF_PREC_FP32, CF_PREC_FP64, ChannelFlow, RunParams, SgsNet,
               z0_of_retau)

FEATURE_NAMES = ("ux", "uy", "uz", "vx", "vy", "vz", "wx", "wy", "wz", "delta", "wall")
STRESS_NAMES = ("t11", "t22", "t33", "t12", "t13", "t23")


def _deal(a):
    """Remove the x and y Nyquist modes of each plane (filter_fftw.f90:316-329)."""
    F = np.fft.rfft2(a)
    ny, nx = a.shape[-2:]
    F[..., :, nx // 2] = 0.0
    F[..., ny // 2, :] = 0.0
    return np.fft.irfft2(F, s=(ny, nx))


def _ddx(a, length):
    F = np.fft.rfft2(a)
    ny, nx = a.shape[-2:]
    k = 2 * math.pi / length * np.arange(nx // 2 + 1)
    k[nx // 2] = 0.0
    return np.fft.irfft2(F * (1j * k), s=(ny, nx))


def _ddy(a, length):
    F = np.fft.rfft2(a)
    ny, nx = a.shape[-2:]
    k = 2 * math.pi / length * np.fft.fftfreq(ny, 1.0 / ny)
    k[ny // 2] = 0.0
    return np.fft.irfft2(F * (1j * k)[:, None], s=(ny, nx))


def coarse_samples(fd, xl, yl, zl):
    """Features (N, 11), targets (N, 6) and the subfilter SGS energy (N,) from
    one filtered snapshot ``fd`` (ChannelFlow.filter_snapshot output)."""
    u = _deal(fd["u"])
    v = _deal(fd["v"])
    wf = _deal(fd["wf"])                      # faces 1..ncz, top = 0
    ncz, ncy, ncx = u.shape
    dz = zl / ncz
    zero = np.zeros((1, ncy, ncx))
    ue = np.concatenate([zero, u, zero])      # levels 0..ncz+1
    ve = np.concatenate([zero, v, zero])
    we = np.concatenate([zero, wf])           # faces 0..ncz
    wxf = _ddx(we, xl)
    wyf = _ddy(we, yl)
    zc = (np.arange(1, ncz + 1) - 0.5) * dz
    wall = np.broadcast_to(np.minimum(zc, zl - zc)[:, None, None], u.shape)
    delta = (xl / ncx * yl / ncy * dz) ** (1.0 / 3.0)
    feats = np.stack([
        _ddx(u, xl), _ddy(u, yl), (ue[2:] - ue[:-2]) / (2 * dz),
        _ddx(v, xl), _ddy(v, yl), (ve[2:] - ve[:-2]) / (2 * dz),
        0.5 * (wxf[:-1] + wxf[1:]), 0.5 * (wyf[:-1] + wyf[1:]), (we[1:] - we[:-1]) / dz,
        np.full(u.shape, delta), wall], axis=-1).reshape(-1, 11)
    ub, vb, wb = fd["u"], fd["v"], fd["wc"]
    tau = np.stack([fd["uu"] - ub * ub, fd["vv"] - vb * vb, fd["ww"] - wb * wb,
                    fd["uv"] - ub * vb, fd["uw"] - ub * wb, fd["vw"] - vb * wb], axis=-1)
    if "s11" in fd:                              # plus the fine grid's own SGS stresses
        tau = tau + np.stack([fd[k] for k in ("s11", "s22", "s33", "s12", "s13", "s23")], axis=-1)
    tau = tau.reshape(-1, 6)
    e_sgs = 0.5 * (tau[:, 0] + tau[:, 1] + tau[:, 2]) + fd["e"].reshape(-1)
    return feats.astype(np.float32), tau.astype(np.float32), e_sgs.astype(np.float32)


def deviatoric(t):
    d = np.array(t, dtype=np.float64, copy=True)
    tr = (d[:, 0] + d[:, 1] + d[:, 2]) / 3.0
    d[:, :3] -= tr[:, None]
    return d


def tke_stress(feats, e_sgs, coef=0.1):
    """-2 K S_ij, K = coef * delta * sqrt(e) (complen.f90:20-21 with alk = dsl)."""
    f = feats.astype(np.float64)
    K = coef * f[:, 9] * np.sqrt(np.maximum(e_sgs.astype(np.float64), 0.0))
    ux, uy, uz, vx, vy, vz, wx, wy, wz = (f[:, i] for i in range(9))
    return np.stack([-2 * K * ux, -2 * K * vy, -2 * K * wz, -K * (uy + vx), -K * (uz + wx),
                     -K * (vz + wy)], axis=-1)


def _dev_err(pred, true):
    """Mean squared deviatoric stress error, off-diagonals counted twice
    (the Frobenius norm of the symmetric tensor)."""
    d = deviatoric(pred) - deviatoric(true)
    w = np.array([1, 1, 1, 2, 2, 2], float)
    return float(np.mean((d * d) @ w))


def evaluate(pred, feats, tau, e_sgs):
    """A-priori comparison on held-out samples; see the module docstring."""
    tke = tke_stress(feats, e_sgs)
    dt, dk = deviatoric(tau), deviatoric(tke)
    w = np.array([1, 1, 1, 2, 2, 2], float)
    c = float(((dt * dk) @ w).sum() / max(((dk * dk) @ w).sum(), 1e-300))
    S = np.stack([feats[:, 0], feats[:, 4], feats[:, 8], 0.5 * (feats[:, 1] + feats[:, 3]),
                  0.5 * (feats[:, 2] + feats[:, 6]), 0.5 * (feats[:, 5] + feats[:, 7])], -1).astype(np.float64)

    def diss(t):
        return -(deviatoric(t) * S) @ w

    eps_true = diss(tau)
    out = {
        "n": int(tau.shape[0]),
        "mse_dev": {"mlp": _dev_err(pred, tau), "tke": _dev_err(tke, tau),
                    "tke_best_coef": _dev_err(c * tke, tau), "zero": _dev_err(0 * tau, tau)},
        "tke_best_coef_scale": c,
        "dissipation_corr": {"mlp": float(np.corrcoef(diss(pred), eps_true)[0, 1]),
                             "tke": float(np.corrcoef(diss(tke), eps_true)[0, 1])},
        "mean_dissipation": {"true": float(eps_true.mean()), "mlp": float(diss(pred).mean()),
                             "tke": float(diss(tke).mean())},
    }
    return out


def _log(*a):
    print(*a, flush=True)


_ENGLEZ, _UXYM = 0, 3                 # meanbuf rows (allocate.f90:89,93)


def _validation_les(args, precision):
    nx, nz = args.val_grid
    z0 = z0_of_retau(args.re_tau)
    dt = args.cfl / (3.2 * (3 * nx / 2) / (2 * math.pi) + 0.5 / (2.0 / nz))
    g = ChannelFlow(RunParams(nx, nx, nz, z0=z0, dt=dt, precision=precision))
    g.seeded_init(args.val_seed, 0.1)
    return g


def _means_after(g, steps):
    """compmn means of the last of ``steps`` RK3 steps."""
    g.step(steps - 1)
    for s in (1, 2, 3):
        g.stage(s, mnout=True)
    g.sync()
    return g.compmn()


def select_candidate(vals, min_accept):
    """Index of the candidate to keep: the lowest validation score among the
    finite-score candidates accepted at a fraction min_accept of the cells or
    more, else among all finite-score ones, else among all."""
    fin = [i for i, v in enumerate(vals) if math.isfinite(v["score"])]
    pool = [i for i in fin if vals[i]["accepted"] >= min_accept] or fin or list(range(len(vals)))
    return min(pool, key=lambda i: vals[i]["score"])


def validation_reference(args):
    """Means of the FP64 TKE-closure run of the validation case."""
    g = _validation_les(args, CF_PREC_FP64)
    try:
        return _means_after(g, args.val_steps)
    finally:
        g.close()


def _validate_on(net, args, ref, grid):
    g = _validation_les(args, CF_PREC_FP32)
    try:
        g.set_closure(net)
        g.set_closure_grid(grid)
        mb = _means_after(g, args.val_steps)
        acc, tot = g.closure_counts()
    finally:
        g.close()
    uxym = float(np.max(np.abs(mb[_UXYM] - ref[_UXYM]) / np.abs(ref[_UXYM])))
    ratio = float(np.sum(mb[_ENGLEZ]) / np.sum(ref[_ENGLEZ]))
    finite = bool(np.all(np.isfinite(mb)))
    score = max(uxym / 0.01, abs(ratio - 1.0) / 0.05) if finite else float("inf")
    return {"uxym": uxym, "englez_ratio": ratio, "accepted": acc / max(tot, 1), "score": score}


def validate(net, args, ref):
    """A-posteriori error of ``net`` on the validation case against the FP64
    TKE-closure means ``ref``, per closure grid: the largest relative error
    of the mean profile uxym, the ratio of the summed resolved TKE englez,
    the accepted fraction and the score max(uxym / 0.01, |ratio - 1| / 0.05).
    The candidate's score is the larger of the two grids' scores, its
    accepted fraction the smaller."""
    out = {"small": _validate_on(net, args, ref, CF_CLOSURE_GRID_SMALL),
           "dealiased": _validate_on(net, args, ref, CF_CLOSURE_GRID_DEALIASED)}
    out["score"] = max(out["small"]["score"], out["dealiased"]["score"])
    out["accepted"] = min(out["small"]["accepted"], out["dealiased"]["accepted"])
    return out


def run(args):
    os.makedirs(args.out_dir, exist_ok=True)
    nx, nz = args.nx, args.nz
    z0 = z0_of_retau(args.re_tau)
    dz = 2.0 / nz
    dt = args.cfl / (3.2 * (3 * nx / 2) / (2 * math.pi) + 0.5 / dz)
    g = ChannelFlow(RunParams(nx, nx, nz, z0=z0, dt=dt, precision=CF_PREC_FP32))
    g.seeded_init(args.seed, 0.1)
    t0 = time.time()
    done = 0
    while done < args.spinup:
        k = min(250, args.spinup - done)
        g.step(k)
        g.sync()
        done += k
        _log(f"[learned] spin-up {done}/{args.spinup} steps, {time.time() - t0:.0f} s, "
             f"utau {g.scalars()['utau']:.4f}")
    filters = [tuple(int(x) for x in f.split(",")) for f in args.filters]
    per = []
    nheld = 0
    held = None
    chunks_f, chunks_t, chunks_e = [], [], []
    for s in range(args.snaps):
        g.step(args.every)
        g.sync()
        for (fx, fy, fz) in filters:
            fd = g.filter_snapshot(fx, fy, fz)
            f, t, e = coarse_samples(fd, 2 * math.pi, math.pi, 2.0)
            if s == args.snaps - 1:
                held = (f, t, e) if held is None else tuple(np.concatenate([a, b]) for a, b in zip(held, (f, t, e)))
                nheld += f.shape[0]
            else:
                chunks_f.append(f); chunks_t.append(t); chunks_e.append(e)
        per.append(time.time() - t0)
        _log(f"[learned] snapshot {s + 1}/{args.snaps}, {time.time() - t0:.0f} s")
    state_stats = {"utau": g.scalars()["utau"], "steps": args.spinup + args.snaps * args.every, "dt": dt}
    g.close()
    ntrain = sum(x.shape[0] for x in chunks_f)
    # standardisation from a subsample of the training set
    sub = np.concatenate([x[:: max(1, x.shape[0] // 20000)] for x in chunks_f])
    subt = np.concatenate([x[:: max(1, x.shape[0] // 20000)] for x in chunks_t])
    in_mu = sub.mean(0); in_sd = sub.std(0)
    in_sd[in_sd < 1e-12 * np.maximum(1.0, np.abs(in_mu))] = 1.0
    out_mu = subt.mean(0); out_sd = subt.std(0)
    net = SgsNet(0)
    net.set_norm(in_mu, in_sd, out_mu, out_sd)
    rng = np.random.default_rng(args.seed)
    net.reserve(ntrain)
    off = 0
    for f, t in zip(chunks_f, chunks_t):
        net.put_samples(off, f, t)
        off += f.shape[0]
    del chunks_f, chunks_t, chunks_e
    _log(f"[learned] {ntrain} training samples on the device, {nheld} held out")
    f, t, e = held
    ref = validation_reference(args)
    cands = []
    for c in range(args.candidates):
        net.set_params(_init_params(np.random.default_rng(args.seed + c)))
        net.reset_optimizer()
        hist = []
        step = 0
        for (nsteps, lr) in args.schedule:
            for k in range(0, nsteps, 500):
                m = min(500, nsteps - k)
                losses = net.train(m, args.batch, lr, seed=args.seed + 1000003 * c + step)
                step += m
                hist.append((step, float(losses[-50:].mean())))
        ev = evaluate(net.forward(f), f, t, e)
        val = validate(net, args, ref)
        cands.append({"params": net.get_params(), "loss_history": hist, "final_loss": hist[-1][1],
                      "held_out": ev, "validation": val})
        _log(f"[learned] candidate {c}: batch loss {hist[-1][1]:.4f}, held-out MSE "
             f"{ev['mse_dev']['mlp']:.3e}, validation {json.dumps(val)}, {time.time() - t0:.0f} s")
    best = select_candidate([cd["validation"] for cd in cands], args.min_accept)
    net.set_params(cands[best]["params"])
    hist, ev = cands[best]["loss_history"], cands[best]["held_out"]
    _log(f"[learned] kept candidate {best}; held-out a-priori:", json.dumps(ev))
    wpath = os.path.join(args.out_dir, "sgs_mlp.cfmlp")
    net.save(wpath)
    meta = {
        "fine_grid": [nx, nx, nz], "re_tau": args.re_tau, "z0": z0, "dt": dt, "cfl": args.cfl,
        "seed": args.seed, "spinup_steps": args.spinup, "snapshots": args.snaps, "every": args.every,
        "filters": filters, "train_samples": ntrain, "held_out_samples": nheld, "batch": args.batch,
        "schedule": args.schedule, "loss_history": hist, "held_out": ev, "run": state_stats,
        "validation_case": {"grid": [args.val_grid[0], args.val_grid[0], args.val_grid[1]],
                            "seed": args.val_seed, "steps": args.val_steps,
                            "reference": "FP64 TKE closure", "closure_grids": ["nnx x nny", "3/2"]},
        "min_accept": args.min_accept,
        "candidates": [{k: cd[k] for k in ("final_loss", "held_out", "validation")} for cd in cands],
        "selected": best,
        "wall_s": time.time() - t0,
    }
    with open(os.path.join(args.out_dir, "sgs_mlp.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    # a small held-out fixture for tests
    idx = rng.choice(f.shape[0], size=min(args.fixture, f.shape[0]), replace=False)
    np.savez(os.path.join(args.out_dir, "sgs_heldout.npz"), feats=f[idx], tau=t[idx], e_sgs=e[idx])
    net.close()
    _log(f"[learned] wrote {wpath}")


def _init_params(rng):
    """He-uniform initial weights (smaller on the output layer), zero biases."""
    layers = ((64, 11), (64, 64), (64, 64), (6, 64))
    out = []
    for (m, n) in layers:
        lim = math.sqrt(6.0 / n) * (0.5 if m == 6 else 1.0)
        out.append(rng.uniform(-lim, lim, (m, n)).ravel())
        out.append(np.zeros(m))
    return np.concatenate(out).astype(np.float32)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--re-tau", type=float, default=395.0)
    ap.add_argument("--cfl", type=float, default=0.3)
    ap.add_argument("--seed", type=int, default=10972)
    ap.add_argument("--spinup", type=int, default=2000)
    ap.add_argument("--snaps", type=int, default=33, help="snapshots; the last one is held out")
    ap.add_argument("--every", type=int, default=25)
    ap.add_argument("--filters", nargs="+", default=["4,4,2", "4,4,4"])
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--schedule", type=lambda s: [(int(a), float(b)) for a, b in
                                                  (p.split(":") for p in s.split(","))],
                    default=[(4000, 1e-3), (3000, 3e-4), (3000, 1e-4)],
                    help="steps:lr,steps:lr,...")
    ap.add_argument("--fixture", type=int, default=16384)
    ap.add_argument("--candidates", type=int, default=8, help="networks trained on the same samples")
    ap.add_argument("--val-grid", type=lambda s: tuple(int(x) for x in s.split(",")), default=(256, 128),
                    help="nnx(=nny),nnz of the a-posteriori validation case")
    ap.add_argument("--val-seed", type=int, default=4242, help="seeded state of the validation case")
    ap.add_argument("--val-steps", type=int, default=200)
    ap.add_argument("--min-accept", type=float, default=0.3,
                    help="least accepted-cell fraction on the validation case for a candidate to be kept")
    ap.add_argument("--out-dir", required=True)
    args = ap.parse_args(argv)
    if args.candidates < 1:
        ap.error("--candidates must be at least 1")
    run(args)


if __name__ == "__main__":
    sys.exit(main())
