"""cf_calc_spectra_noalloc, the allocation-free variant of cf_calc_spectra
(its scratch is a fixed device array, no hipMalloc): through the C-ABI, two
calls in a row on evolving states, the second with fewer levels, each against
cf_calc_spectra on the same state; and the level limit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(a, b):
    ha, va, la = a
    hb, vb, lb = b
    for f in range(3):
        assert np.allclose(ha[f], hb[f], rtol=1e-14, atol=0), f
        assert np.allclose(va[f], vb[f], rtol=1e-14, atol=0), f
    for q in range(6):
        assert np.allclose(la[q], lb[q], rtol=1e-14, atol=0), q


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_spectra_equal_the_allocating_entry_point(precision):
    import channelflow as C
    g = C.ChannelFlow(C.RunParams(64, 64, 80, z0=C.z0_of_retau(180.0), dt=0.002, precision=precision))
    try:
        g.seeded_init(10972, 0.1)
        g.step(2)
        big = list(range(1, 65))                       # 64 levels: the maximum
        _same(g.calc_spectra_noalloc(big), g.calc_spectra(big))
        g.step(1)
        small = [2, 7, 19]                             # a smaller second call on the next state
        _same(g.calc_spectra_noalloc(small), g.calc_spectra(small))
        with pytest.raises(C.ChannelFlowError) as ei:
            g.calc_spectra_noalloc(list(range(1, 66)))    # 65 levels, all valid for nnz = 80
        assert ei.value.code == C.CF_E_ARG
    finally:
        g.close()
