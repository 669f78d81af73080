"""Precision modes against the FP64 path, at the full 256^2 x 128 workload
(Re_tau 395, seeded log-law + noise, dt at CFL 0.3), and the long 32^3 run.

* The FP32 path (FP32 fields, transforms, products and SGS; FP64 pressure,
  projection and update) after 200 steps, against the FP64 run from the same
  state.  The error is printed per output and asserted at about 10x the
  measured value (u 2.0e-7, v 4.4e-5, w 3.5e-5, e 1.1e-5, p 8.8e-4 relative
  L2; mean profile 1.2e-8, resolved TKE 3.3e-6, calc_spectra x-y spectra
  2.4e-5 at every wavenumber).
* The FP32 + learned-closure path after 200 steps, against the FP64
  TKE-closure run.  Asserted: the mean profile within 1 % (measured 0.24 %),
  the plane- and level-summed resolved TKE within 5 % (measured 1.7 %), and
  the x-y spectra at three levels within 120 % per ring and 60 % up to k = 8
  (measured 87 % and 42 %): bounds that a closure with no stress, or with
  stresses off by a factor of a few, fails.  The 5 % spectra target against
  the TKE closure is NOT met.
  The learned closure dissipates about 2.4x less than the TKE closure (its
  a-priori mean SGS dissipation is 0.010 against the TKE's 0.024 and the
  filtered fine run's 0.012).  The seeded noise therefore evolves differently
  near the walls, and the fluctuation fields decorrelate within 200 steps.
* The same with the network on the nnx x nny grid (CF_CLOSURE_GRID_SMALL).
* FP32 against FP64 after 2000 steps at 256^2 x 128 (the metric
  configuration's comparison), per output: fields u 2e-5, e 5e-4, v, w, p
  5e-3; mean profile 2e-5, resolved TKE 3e-4, x-y spectra 3e-3 up to kmax/4
  and 5e-3 over all k (each about 10x the measured error).
* The FP32 path's mean mode: on a horizontally uniform flow the FP32 mean
  profile follows the FP64 path's within 2e-7 over 1000 steps, because the
  plane means of u, v, ru, rv and the imposed gradient are kept in FP64.
* 32^3 for 100 steps (the CPU-reference case): FP64 fields within 1e-9 of
  the oracle.
"""
import math

import numpy as np
import pytest

import cf_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _run(nx, nz, steps, precision, learned=False, small_grid=False):
    import channelflow as C
    z0 = C.z0_of_retau(395.0)
    dt = 0.3 / (3.2 * (3 * nx / 2) / (2 * math.pi) + 0.5 / (2.0 / nz))
    g = C.ChannelFlow(C.RunParams(nx, nx, nz, z0=z0, dt=dt, precision=precision))
    g.seeded_init(10972, 0.1)
    net = None
    if learned:
        import os
        w = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                         "paes-channelflow_amd", "weights", "sgs_mlp.cfmlp")
        net = C.SgsNet(0, w)
        g.set_closure(net)
        if small_grid:
            g.set_closure_grid(C.CF_CLOSURE_GRID_SMALL)
    g.step(steps - 1)
    for s in (1, 2, 3):
        g.stage(s, mnout=True)
    g.sync()
    st = g.get_state()
    mb = g.compmn()
    g.close()
    if net is not None:
        net.close()
    return st, mb


def _spectra(st, nz):
    lv = [nz // 8, nz // 4, nz // 2]
    h, var, _ = O.calc_spectra([(st["u"], st["v"], st["w"])], lv)
    return h, var


def _report(tag, a, b, ma, mb_, nz):
    ha, _ = _spectra(a, nz)
    hb, _ = _spectra(b, nz)
    r = O.MEAN_NAMES.index
    out = {k: rel(a[k], b[k]) for k in "uvwep"}
    out["uxym"] = float(np.max(np.abs(ma[r("uxym")] - mb_[r("uxym")]) / np.abs(mb_[r("uxym")])))
    out["englez"] = rel(ma[r("englez")], mb_[r("englez")])
    out["spectra"] = max(float(np.max(np.abs(x[1:] - y[1:]) / np.abs(y[1:]))) for x, y in zip(ha, hb))
    kmax = ha[0].shape[0] - 1
    for name, lo, hi in (("spectra_k<=8", 1, 8), ("spectra_k<=kmax/4", 1, kmax // 4),
                         ("spectra_k<=kmax/2", 1, kmax // 2)):
        out[name] = max(float(np.max(np.abs(x[lo:hi + 1] - y[lo:hi + 1]) / np.abs(y[lo:hi + 1])))
                        for x, y in zip(ha, hb))
    out["englez_ratio"] = float(np.sum(ma[r("englez")]) / np.sum(mb_[r("englez")]))
    print(tag, {k: f"{v:.2e}" for k, v in out.items()}, flush=True)
    return out


def test_fp32_path_error_per_output_256():
    import channelflow as C
    s64, m64 = _run(256, 128, 200, C.CF_PREC_FP64)
    s32, m32 = _run(256, 128, 200, C.CF_PREC_FP32)
    e = _report("fp32 vs fp64 after 200 steps:", s32, s64, m32, m64, 128)
    bounds = dict(u=2e-6, v=5e-4, w=5e-4, e=1e-4, p=5e-3)   # about 10x the measured errors
    for k in "uvwep":
        assert e[k] < bounds[k], (k, e)
    assert e["uxym"] < 1e-6 and e["englez"] < 5e-5 and e["spectra"] < 3e-4, e


def test_learned_path_against_fp64_tke_256():
    import channelflow as C
    s64, m64 = _run(256, 128, 200, C.CF_PREC_FP64)
    sl, ml = _run(256, 128, 200, C.CF_PREC_FP32, learned=True)
    e = _report("fp32+learned vs fp64 TKE after 200 steps:", sl, s64, ml, m64, 128)
    assert e["uxym"] < 1e-2, e                     # mean profile within 1 %
    assert abs(e["englez_ratio"] - 1.0) < 5e-2, e  # total resolved TKE within 5 %
    # spectra: per ring within 120 % and within 60 % up to k = 8 (measured 87 %
    # and 42 %); the 5 % target is not met (see the module docstring)
    assert e["spectra"] < 1.2 and e["spectra_k<=8"] < 0.6, e


def test_learned_small_grid_against_fp64_tke_256():
    """The network on the nnx x nny grid (CF_CLOSURE_GRID_SMALL), same bounds."""
    import channelflow as C
    s64, m64 = _run(256, 128, 200, C.CF_PREC_FP64)
    sl, ml = _run(256, 128, 200, C.CF_PREC_FP32, learned=True, small_grid=True)
    e = _report("fp32+learned (small grid) vs fp64 TKE after 200 steps:", sl, s64, ml, m64, 128)
    assert e["uxym"] < 1e-2, e
    assert abs(e["englez_ratio"] - 1.0) < 5e-2, e
    assert e["spectra"] < 1.2 and e["spectra_k<=8"] < 0.6, e   # measured 82 % and 37 %


def test_fp32_path_after_2000_steps_256():
    """The metric configuration's own comparison: 2000 steps at 256^2 x 128
    (t = 2.6 at CFL 0.3), FP32 against FP64, both with the TKE closure.
    Measured: mean profile 1.3e-6, resolved TKE 2.4e-5 (rel L2), x-y spectra at
    the reference's levels 2.9e-4 up to kmax/4 and 4.1e-4 over all k; fields u
    2.2e-6, v 2.8e-4, w 3.6e-4, e 4.1e-5, p 3.6e-4 (rel L2)."""
    import channelflow as C
    z0 = C.z0_of_retau(395.0)
    dt = 0.3 / (3.2 * (3 * 256 / 2) / (2 * math.pi) + 0.5 / (2.0 / 128))
    zl = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 18, 24, 30, 36, 48, 64]
    out = {}
    for prec in (C.CF_PREC_FP64, C.CF_PREC_FP32):
        g = C.ChannelFlow(C.RunParams(256, 256, 128, z0=z0, dt=dt, precision=prec))
        g.seeded_init(10972, 0.1)
        for _ in range(4):
            g.step(500)
            g.sync()
            print("precision", prec, "utau", g.scalars()["utau"], flush=True)
        h, _, _ = g.calc_spectra(zl)
        out[prec] = (np.asarray(h), g.compmn(), g.get_state())
        g.close()
    (h32, m32, s32), (h64, m64, s64) = out[C.CF_PREC_FP32], out[C.CF_PREC_FP64]
    fe = {k: rel(s32[k], s64[k]) for k in "uvwep"}
    print("fp32 vs fp64 after 2000 steps, fields:", {k: f"{v:.2e}" for k, v in fe.items()}, flush=True)
    assert fe["u"] < 2e-5 and fe["e"] < 5e-4, fe
    assert fe["v"] < 5e-3 and fe["w"] < 5e-3 and fe["p"] < 5e-3, fe
    r = O.MEAN_NAMES.index
    du = float(np.max(np.abs(m32[r("uxym")] - m64[r("uxym")]) / np.abs(m64[r("uxym")])))
    de = rel(m32[r("englez")], m64[r("englez")])
    kq = (h64.shape[1] - 1) // 4
    ds_q = float(np.max(np.abs(h32[:, 1:kq + 1] - h64[:, 1:kq + 1]) / np.abs(h64[:, 1:kq + 1])))
    ds = float(np.max(np.abs(h32[:, 1:] - h64[:, 1:]) / np.abs(h64[:, 1:])))
    print(f"fp32 vs fp64 after 2000 steps: uxym {du:.2e} englez {de:.2e} spectra k<=kmax/4 {ds_q:.2e} all {ds:.2e}",
          flush=True)
    assert du < 2e-5 and de < 3e-4, (du, de)
    assert ds_q < 3e-3 and ds < 5e-3, (ds_q, ds)


def test_fp32_mean_mode_is_kept_in_fp64():
    """The FP32 path keeps the plane means of u, v and of their RK registers,
    and adds the imposed gradient 0.016 (comp1.f90:152), in FP64
    (press.F90:81-88 pins that mode's pressure).  On a horizontally uniform
    log-law flow only the (0,0) mode evolves: forcing against the wall
    stresses and the SGS z-flux.  Over 1000 steps the FP32 mean profile then
    follows the FP64 path's to the FP32 rounding of its tendencies alone
    (relative 1e-8 per stage at most), where FP32 storage of the profile
    would add half an FP32 ulp (3e-8 relative) at every stage."""
    import channelflow as C
    n, nz = 64, 32
    z0 = C.z0_of_retau(395.0)
    dt = 0.3 / (3.2 * (3 * n / 2) / (2 * math.pi) + 0.5 / (2.0 / nz))
    pr = O.Params(n, n, nz, z0=z0, dt=dt)
    u, v, w, e = O.initial_fields(pr, amp=0.0)
    prof = {}
    for prec in (C.CF_PREC_FP64, C.CF_PREC_FP32):
        g = C.ChannelFlow(C.RunParams(n, n, nz, z0=z0, dt=dt, precision=prec))
        g.set_state(u, v, w, e)
        g.step(1000)
        g.sync()
        prof[prec] = g.compmn()[O.MEAN_NAMES.index("uxym")]
        s = g.get_state()
        prof[(prec, "u")] = s["u"].mean(axis=(1, 2))
        g.close()
    d = float(np.max(np.abs(prof[C.CF_PREC_FP32] - prof[C.CF_PREC_FP64]) / np.abs(prof[C.CF_PREC_FP64])))
    ds = float(np.max(np.abs(prof[(C.CF_PREC_FP32, "u")] - prof[(C.CF_PREC_FP64, "u")])
                      / np.abs(prof[(C.CF_PREC_FP64, "u")])))
    moved = float(np.max(np.abs(prof[C.CF_PREC_FP64] - u.mean(axis=(1, 2))) / np.abs(u.mean(axis=(1, 2)))))
    print(f"uniform flow, 1000 steps: profile change {moved:.2e}; fp32 vs fp64 uxym {d:.2e}, "
          f"downloaded u means {ds:.2e}", flush=True)
    assert moved > 1e-4                        # the profile did evolve
    assert d < 2e-7 and ds < 2e-7, (d, ds)      # measured 2.2e-8


def test_fp64_100_steps_32cubed_matches_oracle():
    import channelflow as C
    pr = O.Params(32, 32, 32, dt=0.01)
    u, v, w, e = O.initial_fields(pr)
    st = O.set_state(pr, u, v, w, e)
    g = C.ChannelFlow(C.RunParams(32, 32, 32, xl=pr.xl, yl=pr.yl, zl=pr.zl, z0=pr.z0, dt=pr.dt,
                                  precision=C.CF_PREC_FP64))
    g.set_state(u, v, w, e)
    errs = []
    for it in range(100):
        O.step(st, pr)
        g.step(1)
        if (it + 1) % 25 == 0:
            gs = g.get_state()
            errs.append(max(rel(gs[k], O.interior(st, pr, k)) for k in "uvwe"))
    print("32^3 FP64 max field error at steps 25/50/75/100:", errs, flush=True)
    assert errs[-1] < 1e-9, errs
    g.close()
