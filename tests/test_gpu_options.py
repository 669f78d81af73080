"""The wall models isurface 1..6 and the Smagorinsky diffusivity (imodel 4)
on the GPU against the FP64 oracle, through the C-ABI.  Two full RK3 steps
(six stages), so case 6 reads a nonzero previous-stage R for its dudt, dvdt.
Both row-transform paths: 32^2 (3/2-grid rows of 48 points shared by the
workgroup) and 128^2 (rows of 192 points owned by wavefronts).  Bound:
relative L2 <= 1e-10 on u, v, w, e after every stage, p <= 1e-10, utau to
1e-12.  The oracle's wall models are pinned against the compiled reference
(tests/test_oracle_ref.py) and by invariants (tests/test_oracle_wall.py).
The FP32 path with each option is held to 1e-4 of the FP64 path at 256^2."""
import numpy as np
import pytest

import cf_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


CASES = ([(s, 3, 32, 16) for s in range(1, 7)] + [(1, 4, 32, 16), (6, 4, 32, 16)]
         + [(s, 3, 128, 16) for s in (2, 3, 4, 5, 6)] + [(3, 4, 128, 16)])


@pytest.mark.parametrize("isurface,imodel,n,nz", CASES)
def test_wall_and_sgs_models_match_oracle(isurface, imodel, n, nz):
    import channelflow as C
    pr = O.Params(n, n, nz, dt=0.005, isurface=isurface, imodel=imodel)
    u, v, w, e = O.initial_fields(pr)
    st = O.set_state(pr, u, v, w, e)
    g = C.ChannelFlow(C.RunParams(n, n, nz, xl=pr.xl, yl=pr.yl, zl=pr.zl, z0=pr.z0, dt=pr.dt,
                                  imodel=imodel, isurface=isurface, precision=C.CF_PREC_FP64))
    g.set_state(u, v, w, e)
    for step in range(2):
        for istage in (1, 2, 3):
            O.stage(st, pr, istage)
            g.stage(istage)
            gs = g.get_state()
            errs = {k: rel(gs[k], O.interior(st, pr, k)) for k in "uvwep"}
            print(isurface, imodel, n, step, istage, {k: f"{x:.1e}" for k, x in errs.items()}, flush=True)
            for k in "uvwe":
                assert errs[k] < 1e-10, (step, istage, k, errs)
            assert errs["p"] < 1e-10, (step, istage, errs)
            assert abs(g.scalars()["utau"] - st.utau) < 1e-12
    g.close()


@pytest.mark.parametrize("isurface,imodel", [(2, 3), (3, 3), (4, 3), (5, 3), (6, 3), (1, 4)])
def test_fp32_options_close_to_fp64(isurface, imodel):
    """The FP32 path with each wall model and with imodel 4, against the FP64
    path (pinned above), at 256^2 (the FP32 row kernel that loads its spectra
    a level ahead): one RK3 step from the same state, within 1e-4 relative
    L2."""
    import channelflow as C
    pr = O.Params(256, 256, 8, dt=0.002, isurface=isurface, imodel=imodel)
    u, v, w, e = O.initial_fields(pr)
    out = {}
    for prec in (C.CF_PREC_FP64, C.CF_PREC_FP32):
        g = C.ChannelFlow(C.RunParams(256, 256, 8, xl=pr.xl, yl=pr.yl, zl=pr.zl, z0=pr.z0, dt=pr.dt,
                                      imodel=imodel, isurface=isurface, precision=prec))
        g.set_state(u, v, w, e)
        g.step(1)
        out[prec] = g.get_state()
        g.close()
    errs = {k: rel(out[C.CF_PREC_FP32][k], out[C.CF_PREC_FP64][k]) for k in "uvwe"}
    print("fp32 vs fp64", isurface, imodel, {k: f"{x:.1e}" for k, x in errs.items()}, flush=True)
    for k in "uvwe":
        assert errs[k] < 1e-4, (k, errs)
