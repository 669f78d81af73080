"""The comp1 clamp of the expanded SGS TKE (comp1.f90:109,121) on the GPU
against the FP64 oracle.  The SGS TKE is set to emin plus sparse sharp
spikes: its band-limited interpolation onto the 3/2 grid undershoots emin
around them, so the clamp fires on many faces (checked on the oracle side
first).  The reference then differentiates the clamped 3/2-grid field
(derivs2.f90:20-21); the device adds the derivatives of the clamp residual
to the e-fluxes (k_eclamp_rows/cols/flux).  One RK3 step, relative L2
<= 1e-10 after every stage, on both row-transform paths (32^2: rows shared
by the workgroup; 128^2: rows owned by wavefronts), with the TKE diffusivity
and with the Smagorinsky one (imodel 4)."""
import numpy as np
import pytest

import cf_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def spiky_e(pr, seed=3):
    rng = np.random.default_rng(seed)
    e = np.full((pr.nnz, pr.nny, pr.nnx), O.EMIN)
    mask = rng.random(e.shape) < 0.02
    e[mask] += rng.uniform(1e-4, 1e-3, mask.sum())
    return e


@pytest.mark.parametrize("n,nz,imodel", [(32, 16, 3), (128, 16, 3), (128, 16, 4)])
def test_clamped_e_matches_oracle(n, nz, imodel):
    import channelflow as C
    pr = O.Params(n, n, nz, dt=0.002, imodel=imodel)
    u, v, w, _ = O.initial_fields(pr)
    e = spiky_e(pr)
    st = O.set_state(pr, u, v, w, e)
    eb = O.expand(O.interior(st, pr, "e"), pr.nnxb, pr.nnyb)
    nclamp = int((eb < O.EMIN).sum())
    print("expanded e below emin at", nclamp, "points", flush=True)
    assert nclamp > 100
    g = C.ChannelFlow(C.RunParams(n, n, nz, xl=pr.xl, yl=pr.yl, zl=pr.zl, z0=pr.z0, dt=pr.dt,
                                  imodel=imodel, precision=C.CF_PREC_FP64))
    g.set_state(u, v, w, e)
    for istage in (1, 2, 3):
        O.stage(st, pr, istage)
        g.stage(istage)
        gs = g.get_state()
        errs = {k: rel(gs[k], O.interior(st, pr, k)) for k in "uvwe"}
        print(n, imodel, istage, {k: f"{x:.1e}" for k, x in errs.items()}, flush=True)
        for k in "uvwe":
            assert errs[k] < 1e-10, (istage, k, errs)
    g.sync()
    g.close()


def test_clamped_e_fp32_runs():
    """The FP32 path on the same state: no error, close to FP64."""
    import channelflow as C
    pr = O.Params(128, 128, 16, dt=0.002)
    u, v, w, _ = O.initial_fields(pr)
    e = spiky_e(pr)
    out = {}
    for prec in (C.CF_PREC_FP64, C.CF_PREC_FP32):
        g = C.ChannelFlow(C.RunParams(128, 128, 16, xl=pr.xl, yl=pr.yl, zl=pr.zl, z0=pr.z0, dt=pr.dt,
                                      precision=prec))
        g.set_state(u, v, w, e)
        g.step(2)
        g.sync()
        out[prec] = g.get_state()
        g.close()
    for k in "uvwe":
        assert rel(out[C.CF_PREC_FP32][k], out[C.CF_PREC_FP64][k]) < 1e-3, k
