#!/usr/bin/env python3
"""Throughput of the channel-flow RK3 timestep on one MI355X.

Metric: grid-point RK3 steps per second (nnx*nny*nnz*steps / time) on the
256x256x128 channel at Re_tau = 395 (z0 = exp(-0.4*5)/395), seeded log-law +
0.1 u_tau noise.  ``value`` is the FP32 path with the learned SGS closure:
- FP32 fields, transforms, products and SGS;
- the BF16-MFMA network at every nnx x nny cell centre (the grid its training
  features live on; --closure-grid dealiased puts it on the 3/2 grid), with
  the TKE closure as fallback;
- FP64 Poisson RHS, tridiagonal solve, projection and update, and FP64 plane
  means of u, v and their RK registers with the imposed gradient.
The network is the trained one shipped in paes-channelflow_amd/weights/.
Timed beside it: the FP64 plain path (TKE closure), the FP32 path with the
TKE closure and the learned path with the network on the other grid.
Fields are resident in HBM before the timed region; the timed region is
exactly K full steps on the library stream bracketed by stream
synchronisation, measured with HIP events.  ``sustained`` then times 500
more steps of the same path, long enough for an outside clock to see.

One JSON line is printed.  ``roofline`` prices the path's dominant kernel
(the one with the longest launch, k_sgs_rows by default): the bytes its
launch must move (its input and output rows, each read or written once) over
its mean launch time from HIP events on the library stream, against the
8 TB/s HBM3E peak; ``roofline.valu`` gives the same kernel against the VALU
issue peak (instructions) and its VALU busy cycles, from counter profiles.
``step_roofline`` prices the whole step at the compulsory bytes of the
persistent arrays (FP32 path: 85 FP32 + 12 FP64 array passes = 436 B per
grid-point step; FP64 path: 97 x 8 = 776 B).  ``cpu_baseline`` times the
FP64 numpy oracle (the CPU restatement of the reference stage) on one whole
RK3 step of the same 256x256x128 workload, its DFTs threaded over the host
cores this process may use.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "paes-channelflow_amd"))
WEIGHTS = os.path.join(ROOT, "paes-channelflow_amd", "weights", "sgs_mlp.cfmlp")

HBM_PEAK_GBS = 8000.0
BYTES_PER_GPSTEP = {"fp32": 436.0, "fp64": 776.0}


def kernel_bytes(name, n, s, istage=2, lean=False):
    """Bytes one launch of kernel ``name`` must read and write, for N = n grid
    points and s-byte reals (complex = 2s; band spectra and x-spectral rows of
    the nnx x nny plane hold N*s bytes per field, 3/2-grid rows 1.5*N*s).
    ``lean``: the 3/2-grid sweep of the nnx x nny closure mode (no velocity
    gradients)."""
    return {
        # U, V, W band spectra + x-spectral e in; u,uy,v,vy,w,wy,e,ey rows out
        # (lean: u, v, w, e, ey)
        "k_col_synth": 4 * n * s + (7.5 if lean else 12) * n * s,
        # U, V, W of the level and W of the level below in (band spectra)
        "k_divz_band": 4 * n * s,
        # 8 synthesis rows in, 8 flux rows (Gu,Fyu,Gv,Fyv,Gw,Fyw,Ge,Fye) out
        # (lean: u, v, w, e, ey in)
        "k_rows": (7.5 if lean else 12) * n * s + 12 * n * s,
        # 8 flux rows in; ru, rv, rw band spectra + x-spectral re out
        "k_col_anal": 12 * n * s + 4 * n * s,
        # U,V,W, R_old, R_new in; U,V,W, R out; FP64 pressure spectrum out
        "k_press": 15 * n * s + 8 * n,
        # x-spectral re, e, re in; e, re, x-spectral e out
        "k_erows": (6 if istage > 1 else 5) * n * s,
        # small-grid closure: U, V, W band spectra in; u, v, w, uy, vy, wy small
        # x-spectral rows out
        "k_sgs_cols_in": 9 * n * s,
        # 6 small rows + physical e in; G and F^y rows of u, v, w out
        "k_sgs_rows": 13 * n * s,
        # G and F^y rows of u, v, w in; ru, rv, rw read and written
        "k_sgs_cols_out": 12 * n * s,
    }.get(name)

BASIS = {
    "k_rows": "24 x N x 4 B per launch (3/2-grid x-spectral rows in and out, once each)",
    "k_rows_mlp": "24 x N x 4 B per launch (3/2-grid x-spectral rows in and out, once each; "
                  "the network's work is on-chip)",
    "k_sgs_rows": "13 x N x 4 B per launch (6 small x-spectral rows and the physical e in, "
                  "6 divergence rows out; the network's work is on-chip)",
    "k_rows_lean": "19.5 x N x 4 B per launch (the u, v, w, e, ey 3/2-grid rows in, the 8 flux "
                   "rows out, once each)",
}

# HBM bytes per dispatch from the FETCH_SIZE and WRITE_SIZE counters (KiB per
# dispatch x 1024) for the 256x256x128 FP32 workload.  FETCH_SIZE counts
# contiguous full-line reads at half their size (a 64 MiB contiguous read by
# k_cast reads back as 32 MiB), so k_rows' row reads take a factor 2;
# WRITE_SIZE is 1:1 (a 256 MiB fill reads back as 256 MiB).
# profiles/k_rows_pmc_256x256x128.md holds the runs.
MEASURED_TRAFFIC = {
    ("k_rows", 256, 128, "fp32"): round((2 * 198182.1 + 391680.0) * 1024),
    # profiles/k_rows_mlp_counters_256x256x128.md
    ("k_rows_mlp", 256, 128, "fp32"): round((2 * 202212.8 + 392256.0) * 1024),
    # profiles/learned_small_grid_counters_256x256x128.md
    ("k_sgs_rows", 256, 128, "fp32"): round((2 * 117372.5 + 196736.0) * 1024),
    ("k_rows_lean", 256, 128, "fp32"): round((2 * 127828.4 + 414662.1) * 1024),
}

# SQ_INSTS_VALU per dispatch (wavefront instructions) at 256x256x128 FP32, from
# the same counter profiles.  The VALU issue peak: 256 CUs x 4 SIMDs x 2.4 GHz,
# one wave64 VALU instruction per 4 cycles per SIMD (16 lanes; no packed FP32).
MEASURED_VALU = {
    ("k_rows", 256, 128, "fp32"): 227.93e6,
    ("k_rows_mlp", 256, 128, "fp32"): 738.28e6,
    ("k_sgs_rows", 256, 128, "fp32"): 290.79e6,
    ("k_rows_lean", 256, 128, "fp32"): 168.49e6,
}
# SQ_ACTIVE_INST_VALU per dispatch (4-cycle issue slots the waves spent in
# VALU instructions; a transcendental takes two) where profiled
# (profiles/learned_small_grid_counters_256x256x128.md).
MEASURED_VALU_ACTIVE = {
    ("k_sgs_rows", 256, 128, "fp32"): 343.09e6,
    ("k_rows_lean", 256, 128, "fp32"): 169.66e6,
}
VALU_PEAK = 256 * 4 * 2.4e9 / 4
# SQ_INSTS_MFMA per dispatch (v_mfma_f32_32x32x16_bf16, 32 x 32 x 16 x 2 flop
# each) from the same counter profiles, against the dense BF16 peak.
MEASURED_MFMA = {
    ("k_sgs_rows", 256, 128, "fp32"): 5.95e6,
}
MFMA_FLOP = 32 * 32 * 16 * 2
BF16_PEAK_TFLOPS = 2500.0


def valu_roofline(kernel, args, ms):
    """The dominant kernel against the VALU issue peak: its SQ_INSTS_VALU per
    launch (counter profile) / its launch time.  None where not profiled."""
    insts = MEASURED_VALU.get((kernel, args.nx, args.nz, "fp32"))
    if insts is None:
        return None
    achieved = insts / (ms * 1e-3)
    out = {"achieved": achieved, "peak": VALU_PEAK, "unit": "wave64 VALU instructions/s",
           "frac": achieved / VALU_PEAK, "insts_per_launch": insts}
    active = MEASURED_VALU_ACTIVE.get((kernel, args.nx, args.nz, "fp32"))
    if active is not None:
        out["busy_frac"] = active / (ms * 1e-3) / VALU_PEAK
    return out


def mfma_roofline(kernel, args, ms):
    """The dominant kernel's BF16 matrix-core work against the dense BF16
    peak: SQ_INSTS_MFMA per launch x flop per instruction / launch time.
    None where not profiled."""
    n = MEASURED_MFMA.get((kernel, args.nx, args.nz, "fp32"))
    if n is None:
        return None
    achieved = n * MFMA_FLOP / (ms * 1e-3) / 1e12
    return {"achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / BF16_PEAK_TFLOPS, "mfma_per_launch": n}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--re-tau", type=float, default=395.0)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 plain path")
    ap.add_argument("--no-tke32", action="store_true", help="skip the FP32 TKE-closure path")
    ap.add_argument("--weights", default=WEIGHTS, help="learned closure network file")
    ap.add_argument("--closure-grid", choices=("dealiased", "small"), default="small",
                    help="evaluate the network on the nnx x nny grid (default) or on the 3/2 grid")
    ap.add_argument("--no-other-grid", action="store_true",
                    help="skip the learned path with the network on the other grid")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-cores", type=int, default=0,
                    help="threads of the CPU baseline's DFTs (default: the cores this process may use)")
    ap.add_argument("--cpu-sample-nz", type=int, default=0,
                    help="levels of the CPU-baseline sample (default: the full nz)")
    ap.add_argument("--sustain", type=int, default=500,
                    help="steps of the headline path timed once more after the K steps (0: none)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one device copy of 256 MiB (counter calibration under rocprofv3)")
    return ap.parse_args()


def run_gpu(args, precision_name, learned=False, grid=None, sustain=0):
    import numpy as np
    import channelflow as C
    nx, nz = args.nx, args.nz
    z0 = C.z0_of_retau(args.re_tau)
    dz = 2.0 / nz
    # fixed dt at CFL ~ 0.3 of the 3/2 grid for |u| ~ 3 u_tau/kappa-scale velocities
    dt = 0.3 / (3.2 * (3 * nx / 2) / (2 * math.pi) + 0.5 / dz)
    prec = C.CF_PREC_FP32 if precision_name == "fp32" else C.CF_PREC_FP64
    rp = C.RunParams(nx, nx, nz, z0=z0, dt=dt, precision=prec)
    g = C.ChannelFlow(rp)
    u, v, w, e = seeded_fields(nx, nz, z0)
    g.set_state(u, v, w, e)
    net = None
    if learned:
        net = C.SgsNet(0, args.weights)
        g.set_closure(net)
        g.set_closure_grid(C.CF_CLOSURE_GRID_SMALL if (grid or args.closure_grid) == "small"
                           else C.CF_CLOSURE_GRID_DEALIASED)
    if args.warmup > 0:
        g.step(args.warmup)
    g.sync()
    t0 = time.perf_counter()
    ms = g.time_steps(args.steps)
    wall = time.perf_counter() - t0
    sustained = None
    if sustain:
        t1 = time.perf_counter()
        ms_s = g.time_steps(sustain)
        sustained = {"steps": sustain, "ms_per_step": ms_s / sustain,
                     "host_wall_s": time.perf_counter() - t1}
    _, per_kernel = g.time_steps(1, per_kernel=True)
    g.sync()
    st = g.get_state()
    finite = all(bool(np.isfinite(st[k]).all()) for k in "uvwe")
    acc = g.closure_counts() if learned else (0, 0)
    out = dict(ms=ms, wall=wall, per_kernel=per_kernel, dt=dt, z0=z0, finite=finite,
               device_bytes=g.device_bytes(), accepted=acc, sustained=sustained)
    g.close()
    if net is not None:
        net.close()
    return out


def seeded_fields(nx, nz, z0):
    """The seeded log-law + noise initial condition (the oracle's recipe,
    computed here directly so bench does not import the oracle)."""
    import numpy as np
    ut = math.sqrt(0.016 * 1.0)
    zl = 2.0
    dz = zl / nz
    zm = (np.arange(1, nz + 1) - 0.5) * dz
    zw = np.minimum(zm, zl - zm)
    prof = ut / 0.4 * np.log(zw / z0)
    rng = np.random.default_rng(10972)
    shp = (nz, nx, nx)
    u = prof[:, None, None] + 0.1 * ut * rng.uniform(-1, 1, shp)
    v = 0.1 * ut * rng.uniform(-1, 1, shp)
    w = 0.1 * ut * rng.uniform(-1, 1, shp)
    w[-1] = 0.0
    e = np.full(shp, 0.01 * ut * ut)
    return u, v, w, e


def cpu_cores():
    """Host cores this process may use (its affinity set, capped by OMP_NUM_THREADS when set)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(omp))) if omp.isdigit() else max(1, n)


def cpu_baseline(args):
    """FP64 numpy oracle (the CPU restatement) on one whole RK3 step (three
    stages) of the same grid, reported in grid-point RK3 steps/s = points /
    step time.  Its DFTs (about 57 % of the step on one core) run on the
    host's cores through scipy.fft's threads; the rest of the step is
    single-threaded numpy."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cf_oracle as O
    cores = args.cpu_cores or cpu_cores()
    nz = args.cpu_sample_nz or args.nz
    pr = O.Params(args.nx, args.nx, nz, z0=math.exp(-2.0) / args.re_tau, dt=0.002)
    st = O.set_state(pr, *O.initial_fields(pr))
    O.FFT_WORKERS = cores
    try:
        t0 = time.perf_counter()
        O.step(st, pr)
        t = time.perf_counter() - t0
    finally:
        O.FFT_WORKERS = 1
    pts = args.nx * args.nx * nz
    return {"value": pts / t, "unit": "grid-pt RK3 steps/s", "cores": cores, "kind": "port",
            "sample": f"one RK3 step (three stages) of the FP64 numpy oracle at "
                      f"{args.nx}x{args.nx}x{nz}, {t:.1f} s; DFTs on {cores} threads "
                      f"(scipy.fft), the rest single-threaded"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the timestep runs on one GPU; --gpus must be 1 "
                                   "(the multi-GPU z-slab decomposition is designed, not built)"}))
        sys.exit(2)
    pts = args.nx * args.nx * args.nz
    if args.copy_probe:
        import torch
        src = torch.ones(64 * 2 ** 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print("[bench] copy probe: 256 MiB read + 256 MiB written", file=sys.stderr, flush=True)
    print(f"[bench] fp32 + learned closure {args.nx}x{args.nx}x{args.nz}", file=sys.stderr, flush=True)
    r32 = run_gpu(args, "fp32", learned=True, sustain=args.sustain)
    value = pts * args.steps / (r32["ms"] * 1e-3)
    other = "dealiased" if args.closure_grid == "small" else "small"
    ro = None
    if not args.no_other_grid:
        print(f"[bench] fp32 + learned closure, network on the {other} grid", file=sys.stderr, flush=True)
        ro = run_gpu(args, "fp32", learned=True, grid=other)
    print("[bench] fp64 plain path", file=sys.stderr, flush=True)
    r64 = None if args.no_fp64 else run_gpu(args, "fp64")
    print("[bench] fp32 TKE path", file=sys.stderr, flush=True)
    rt = None if args.no_tke32 else run_gpu(args, "fp32")
    kern = r32["per_kernel"]
    dom = max((k for k in kern if not k.startswith("(")), key=lambda k: kern[k])
    step_ms = r32["ms"] / args.steps
    step_achieved = BYTES_PER_GPSTEP["fp32"] * value / 1e9
    lean = args.closure_grid == "small"
    dom_bytes = kernel_bytes(dom, pts, 4, lean=lean)
    achieved = dom_bytes / (kern[dom] * 1e-3) / 1e9
    per_kernel_gbs = {k: kernel_bytes(k, pts, 4, lean=lean) / (t * 1e-3) / 1e9
                      for k, t in kern.items() if kernel_bytes(k, pts, 4, lean=lean) and t > 0}
    # on the 3/2 grid the learned path runs the row kernel with the network
    # inside (k_rows_mlp): same row inputs and outputs, same algorithmic bytes;
    # on the nnx x nny grid it runs the lean row sweep (k_rows<T, NX, true>)
    dom_label = dom
    if dom == "k_rows":
        dom_label = "k_rows_lean" if lean else "k_rows_mlp"
    traffic = MEASURED_TRAFFIC.get((dom_label, args.nx, args.nz, "fp32"))
    line = {
        "metric": "grid-pt RK3 steps/s @256^2x128: FP32+learned SGS vs FP64 plain; % HBM peak",
        "value": value,
        "unit": "grid-pt RK3 steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (fp64 pressure/projection; bf16 MFMA closure)",
        "data": "synthetic: seeded log-law + 0.1 u_tau uniform noise, e = 0.01 u_tau^2",
        "config": {
            "workload": f"channel {args.nx}x{args.nx}x{args.nz}, 2pi x pi x 2, Re_tau={args.re_tau:g} "
                        f"(z0={r32['z0']:.4e}), learned SGS closure (MLP 11-64-64-64-6, BF16 MFMA, on the "
                        f"{'nnx x nny' if args.closure_grid == 'small' else '3/2'} grid) "
                        f"with TKE fallback (imodel 3), Schumann wall (isurface 1), "
                        f"3/2 dealiasing, dt={r32['dt']:.3e}",
            "grid": [args.nx, args.nx, args.nz],
            "precision": "fp32 fields/FFT/SGS, fp64 pressure, bf16 MFMA network with fp32 accumulation",
            "learned_sgs": True,
            "closure_grid": args.closure_grid,
            "closure_cells_accepted": r32["accepted"][0] / max(r32["accepted"][1], 1),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": dom_label,
            "bytes_per_launch": dom_bytes,
            "ms_per_launch": kern[dom],
            "basis": f"{dom_label}: {BASIS.get(dom_label, 'algorithmic bytes per launch')} "
                     "/ mean launch time from HIP events on the library stream",
            "valu": valu_roofline(dom_label, args, kern[dom]),
            "mfma": mfma_roofline(dom_label, args, kern[dom]),
        },
        "step_roofline": {
            "achieved": step_achieved,
            "frac": step_achieved / HBM_PEAK_GBS,
            "basis": "436 compulsory HBM bytes per grid-point RK3 step x grid-pt steps/s",
        },
        "sustained": r32["sustained"],
        "device_bytes": r32["device_bytes"],
        "kernel_ms_per_stage": kern,
        "kernel_gbs": per_kernel_gbs,
        "fp64_plain": None,
        "fp32_tke": None,
        "finite": r32["finite"],
    }
    if rt is not None:
        vt = pts * args.steps / (rt["ms"] * 1e-3)
        line["fp32_tke"] = {"value": vt, "ms_per_step": rt["ms"] / args.steps,
                            "kernel_ms_per_stage": rt["per_kernel"], "finite": rt["finite"],
                            "device_bytes": rt["device_bytes"]}
    if ro is not None:
        line["learned_other_grid"] = {
            "closure_grid": other, "value": pts * args.steps / (ro["ms"] * 1e-3),
            "ms_per_step": ro["ms"] / args.steps, "kernel_ms_per_stage": ro["per_kernel"],
            "closure_cells_accepted": ro["accepted"][0] / max(ro["accepted"][1], 1), "finite": ro["finite"]}
    if r64 is not None:
        v64 = pts * args.steps / (r64["ms"] * 1e-3)
        line["fp64_plain"] = {
            "value": v64, "ms_per_step": r64["ms"] / args.steps,
            "hbm_frac": BYTES_PER_GPSTEP["fp64"] * v64 / 1e9 / HBM_PEAK_GBS,
            "learned_fp32_over_fp64": value / v64, "kernel_ms_per_stage": r64["per_kernel"],
            "finite": r64["finite"], "device_bytes": r64["device_bytes"],
        }
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (numpy oracle, one step)", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(args)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
