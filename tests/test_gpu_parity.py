"""GPU path vs the FP64 oracle on the same seeded fields, through the C-ABI
(ctypes).  Grid sizes cover both row-transform paths: nnx = 32, 64 (3/2-grid
rows of 48, 96 points shared by the workgroup) and nnx = 128, 256, 512 (rows
of 192, 384, 768 points owned by single wavefronts, the path of the large
workloads).  Tolerances: relative L2 <= 1e-10 for every field of the FP64
path, the pressure included; the FP32 path (FP64 pressure) is checked at
2e-5 (32^3) and 5e-5 (256^2) after one step."""
import numpy as np
import pytest

import cf_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def make(pr, precision):
    import channelflow as C
    rp = C.RunParams(pr.nnx, pr.nny, pr.nnz, xl=pr.xl, yl=pr.yl, zl=pr.zl, z0=pr.z0, dt=pr.dt,
                     precision=precision)
    return C.ChannelFlow(rp)


def fields(st, pr):
    return {k: O.interior(st, pr, k) for k in "uvwep"}


@pytest.mark.parametrize("n,nz", [(32, 32), (64, 16), (128, 24), (256, 10), (512, 4)])
def test_fp64_stages_match_oracle(n, nz):
    import channelflow as C
    pr = O.Params(n, n, nz, dt=0.01)
    u, v, w, e = O.initial_fields(pr)
    st = O.set_state(pr, u, v, w, e)
    g = make(pr, C.CF_PREC_FP64)
    g.set_state(u, v, w, e)
    s0 = g.get_state()
    for k in "uvwe":
        assert rel(s0[k], O.interior(st, pr, k)) < 1e-13, k
    for istage in (1, 2, 3):
        O.stage(st, pr, istage)
        g.stage(istage)
        gs = g.get_state()
        ref = fields(st, pr)
        errs = {k: rel(gs[k], ref[k]) for k in "uvwep"}
        print(n, nz, istage, errs, flush=True)
        for k in "uvwe":
            assert errs[k] < 1e-10, (istage, k, errs)
        assert errs["p"] < 1e-10, (istage, errs)
        assert abs(g.scalars()["utau"] - st.utau) < 1e-12
        dz = g.divz()
        assert np.allclose(dz, st.divz, rtol=1e-8, atol=1e-25)
    g.close()


@pytest.mark.parametrize("nz,xl,kernel", [
    (9, 2 * np.pi, "partitioned: chunks of 4, 4 and 1 levels"),
    (13, 2 * np.pi, "partitioned: chunks of 4, 4, 4 and 1 levels"),
    (129, 2 * np.pi, "serial: 33 chunks of 4 levels exceed 32"),
    (16, 32.0, "serial: k^2 = (2 pi / 32)^2 < 0.05"),
])
def test_fp64_pressure_kernels_match_oracle(nz, xl, kernel):
    """The pressure solve, projection and update per mode: the partitioned
    kernel (z-chunks of 4 levels on the FP64 path, a tridiagonal interface
    system) at chunkings whose last chunk holds one level, and the serial
    sweep where the library selects it (more than 32 chunks, or a domain long
    enough that k^2 < 0.05), all against the oracle's serial tridv after a
    full step."""
    import channelflow as C
    pr = O.Params(32, 32, nz, xl=xl, dt=0.004)
    u, v, w, e = O.initial_fields(pr)
    st = O.set_state(pr, u, v, w, e)
    g = make(pr, C.CF_PREC_FP64)
    g.set_state(u, v, w, e)
    O.step(st, pr)
    g.step(1)
    gs = g.get_state()
    ref = fields(st, pr)
    errs = {k: rel(gs[k], ref[k]) for k in "uvwep"}
    print(kernel, errs, flush=True)
    for k in "uvwep":
        assert errs[k] < 1e-10, (kernel, k, errs)
    g.close()


@pytest.mark.parametrize("nz", [17, 40])
def test_fp32_pressure_chunkings_close_to_fp64(nz):
    """FP32 path (z-chunks of 8 levels: 8, 8, 1 at nz = 17; 5 full chunks at
    40) after one step against the FP64 path on the same state."""
    import channelflow as C
    pr = O.Params(64, 64, nz, dt=0.004)
    u, v, w, e = O.initial_fields(pr)
    outs = []
    for prec in (C.CF_PREC_FP32, C.CF_PREC_FP64):
        g = make(pr, prec)
        g.set_state(u, v, w, e)
        g.step(1)
        outs.append(g.get_state())
        g.close()
    errs = {k: rel(outs[0][k], outs[1][k]) for k in "uvwep"}
    print(nz, errs, flush=True)
    for k in "uvwe":
        assert errs[k] < 5e-5, (k, errs)
    assert errs["p"] < 5e-3, errs


def test_fp64_tendencies_match_oracle():
    import channelflow as C
    pr = O.Params(32, 32, 32, dt=0.01)
    u, v, w, e = O.initial_fields(pr)
    st = O.set_state(pr, u, v, w, e)
    g = make(pr, C.CF_PREC_FP64)
    g.set_state(u, v, w, e)
    for istage in (1, 2):
        # comp1 output before the projection
        ref = st.copy()
        dtab2 = pr.dt * O.ZETAS[istage - 1]
        O.lower(ref, pr); O.upper(ref, pr)
        O.comp1(ref, pr, istage, False, dtab2)
        O.stage(st, pr, istage)
        g.stage(istage)
        t1 = g.get_comp1_tendency()
        for k in ("ru", "rv", "rw"):
            assert rel(t1[k], O.interior(ref, pr, k)) < 1e-10, (istage, k)
        tt = g.get_tendency()
        for k in ("ru", "rv", "rw", "re"):
            assert rel(tt[k], O.interior(st, pr, k)) < 1e-10, (istage, k)
    g.close()


def test_fp64_ten_steps_and_means():
    import channelflow as C
    pr = O.Params(32, 32, 32, dt=0.01)
    u, v, w, e = O.initial_fields(pr)
    st = O.set_state(pr, u, v, w, e)
    g = make(pr, C.CF_PREC_FP64)
    g.set_state(u, v, w, e)
    for it in range(10):
        mn = (it == 9)
        for s in (1, 2, 3):
            O.stage(st, pr, s, mnout=mn)
            g.stage(s, mnout=mn)
    gs = g.get_state()
    ref = fields(st, pr)
    for k in "uvwe":
        assert rel(gs[k], ref[k]) < 1e-10, k
    mb = g.compmn()
    mo = O.compmn(st, pr)
    for row in range(19):
        if row == 12:
            continue
        # wxym is round-off (|w| ~ 1e-2): compare it at an absolute 1e-16
        atol = 1e-16 if O.MEAN_NAMES[row] == "wxym" else 1e-22
        assert np.allclose(mb[row], mo[row], rtol=1e-8, atol=atol), O.MEAN_NAMES[row]
    g.close()


def test_fp32_step_close_to_oracle():
    import channelflow as C
    pr = O.Params(32, 32, 32, dt=0.01)
    u, v, w, e = O.initial_fields(pr)
    st = O.set_state(pr, u, v, w, e)
    g = make(pr, C.CF_PREC_FP32)
    g.set_state(u, v, w, e)
    O.step(st, pr)
    g.step(1)
    gs = g.get_state()
    ref = fields(st, pr)
    for k in "uvwe":
        assert rel(gs[k], ref[k]) < 2e-5, (k, rel(gs[k], ref[k]))
    g.close()


def test_fp32_wave_path_step_close_to_oracle():
    import channelflow as C
    pr = O.Params(256, 256, 10, dt=0.002)
    u, v, w, e = O.initial_fields(pr)
    st = O.set_state(pr, u, v, w, e)
    g = make(pr, C.CF_PREC_FP32)
    g.set_state(u, v, w, e)
    O.step(st, pr)
    g.step(1)
    gs = g.get_state()
    ref = fields(st, pr)
    for k in "uvwe":
        assert rel(gs[k], ref[k]) < 5e-5, (k, rel(gs[k], ref[k]))
    g.close()


@pytest.mark.parametrize("n,nz", [(128, 16), (512, 4)])
def test_fp32_step_close_to_fp64_step(n, nz):
    """The FP32 TKE path against the FP64 path, which the oracle pins, from the
    same state, at the sizes whose FP32 kernels no other test runs: 128^2
    (three-wavefront rows) and 512^2 (twelve-wavefront rows, eight-column
    column tiles).  One RK3 step, within 5e-5 relative L2, as at 256^2."""
    import channelflow as C
    pr = O.Params(n, n, nz, dt=0.002)
    u, v, w, e = O.initial_fields(pr)
    out = {}
    for prec in (C.CF_PREC_FP64, C.CF_PREC_FP32):
        g = make(pr, prec)
        g.set_state(u, v, w, e)
        g.step(1)
        out[prec] = g.get_state()
        g.close()
    for k in "uvwe":
        err = rel(out[C.CF_PREC_FP32][k], out[C.CF_PREC_FP64][k])
        assert err < 5e-5, (n, k, err)


def test_fp64_128x64_fields_divergence_means_spectra():
    """The 128^2 x 64 channel (Re_tau 180, dt = 1/256): three steps on the GPU
    and in the oracle.  Fields within 1e-10, the small-grid divergence at the
    oracle's round-off level, the compmn mean profile and resolved TKE, and
    the calc_spectra x-y spectra of the downloaded fields."""
    import channelflow as C
    pr = O.Params(128, 128, 64, dt=1.0 / 256)
    u, v, w, e = O.initial_fields(pr)
    st = O.set_state(pr, u, v, w, e)
    g = make(pr, C.CF_PREC_FP64)
    g.set_state(u, v, w, e)
    for it in range(3):
        mn = (it == 2)
        for s in (1, 2, 3):
            O.stage(st, pr, s, mnout=mn)
            g.stage(s, mnout=mn)
    gs = g.get_state()
    ref = fields(st, pr)
    for k in "uvwe":
        assert rel(gs[k], ref[k]) < 1e-10, (k, rel(gs[k], ref[k]))
    # divergence of the GPU fields, measured by the oracle's small-grid operator
    gst = O.set_state(pr, gs["u"], gs["v"], gs["w"], gs["e"])
    dg = O.divergence(gst, pr)
    do = O.divergence(st, pr)
    scale = np.abs(O.xderiv(ref["u"], pr.xl)).max()
    assert np.abs(dg).max() < 1e-11 * scale and np.abs(do).max() < 1e-11 * scale
    mb, mo = g.compmn(), O.compmn(st, pr)
    for name in ("uxym", "englez", "engsbz", "uule", "vvle", "wwle", "uwle"):
        r = O.MEAN_NAMES.index(name)
        assert np.allclose(mb[r], mo[r], rtol=1e-8, atol=1e-20), name
    lv = [8, 16, 32, 48]
    hg, vg, _ = O.calc_spectra([(gs["u"], gs["v"], gs["w"])], lv)
    ho, vo, _ = O.calc_spectra([(ref["u"], ref["v"], ref["w"])], lv)
    for f in range(3):
        assert np.allclose(hg[f], ho[f], rtol=1e-8, atol=1e-30), f
        assert np.allclose(vg[f], vo[f], rtol=1e-8), f
    # calc_spectra on the device, at the reference's own level list (calc_spectra.f90:32)
    zl = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 18, 24, 30, 36, 48]
    hd, vd, ld = g.calc_spectra(zl)
    ho, vo, lo = O.calc_spectra([(ref["u"], ref["v"], ref["w"])], zl)
    for f in range(3):
        scale = np.abs(ho[f][1:]).max()
        assert np.allclose(hd[f][1:], ho[f][1:], rtol=1e-8, atol=1e-12 * scale), f
        assert np.allclose(hd[f][0], ho[f][0], rtol=1e-6, atol=1e-20), f
        assert np.allclose(vd[f], vo[f], rtol=1e-8), f
    for q in range(6):
        assert np.allclose(ld[q], lo[q], rtol=1e-8), q
    g.close()
