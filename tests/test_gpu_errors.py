"""Device-side error reporting through the C-ABI: a non-finite SGS TKE
(CF_E_NONFINITE) is raised by the next cf_sync.  The NaN is planted in the
host input of cf_set_state, so every kernel runs its ordinary arithmetic on
it and exits normally; the e update (comp2.f90:38) tests the updated e for
finiteness before the clamp e = max(e, emin), which would otherwise turn a
NaN into emin."""
import numpy as np
import pytest

import cf_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("field", ["e", "u"])
@pytest.mark.parametrize("precision", [0, 1])
def test_nonfinite_tke_is_reported(precision, field):
    """A NaN in e is caught by the e update of the same stage; a NaN in u
    reaches e in the same stage through the TKE advection."""
    import channelflow as C
    pr = O.Params(32, 32, 16, dt=0.01)
    u, v, w, e = O.initial_fields(pr)
    (e if field == "e" else u)[8, 5, 7] = np.nan   # one interior level, away from both walls
    g = C.ChannelFlow(C.RunParams(32, 32, 16, z0=pr.z0, dt=pr.dt, precision=precision))
    try:
        g.set_state(u, v, w, e)
        g.stage(1)
        with pytest.raises(C.ChannelFlowError) as ei:
            g.sync()
        assert ei.value.code == C.CF_E_NONFINITE, ei.value
    finally:
        g.close()


def test_finite_run_reports_no_error():
    import channelflow as C
    pr = O.Params(32, 32, 16, dt=0.01)
    u, v, w, e = O.initial_fields(pr)
    g = C.ChannelFlow(C.RunParams(32, 32, 16, z0=pr.z0, dt=pr.dt, precision=C.CF_PREC_FP32))
    try:
        g.set_state(u, v, w, e)
        g.step(3)
        g.sync()
    finally:
        g.close()
