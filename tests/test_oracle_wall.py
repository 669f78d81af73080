"""Invariants of the oracle's wall models (cf_oracle.wall_stress) that the
compiled reference cannot pin.  Case 2 (Moeng) reads fluctuation fields the
reference leaves unset for it, so it is pinned here:

* For a uniform wall-level plane, Moeng's model (2) and its unlinearised form
  (4) both reduce to Schumann's (1): wind = <wind> = vsfc and the
  fluctuations vanish.
* Case 2 is case 4 linearised in the fluctuations: with the fluctuations
  scaled by eps their difference is O(eps^2) relative to the stress, so it
  falls 100x when eps falls 10x.  A dropped or sign-flipped fluctuation term
  in either case leaves an O(eps) difference and fails.
Both walls, with ugal = vgal = 0 (the reference's channel setting).
"""
import numpy as np
import pytest

import cf_oracle as O


def _stress(isurface, ub, vb, upper):
    pr = O.Params(16, 16, 8, dt=0.01, isurface=isurface)
    st = O.State.__new__(O.State)
    st.utau_s = st.dudt = st.dvdt = 0.0
    tuw, tvw, _, _ = O.wall_stress(pr, ub, vb, upper, st)
    return np.asarray(tuw, float), np.asarray(tvw, float)


@pytest.mark.parametrize("upper", [False, True])
def test_uniform_plane_reduces_to_schumann(upper):
    ub = np.full((24, 24), 1.7)
    vb = np.full((24, 24), -0.4)
    t1 = _stress(1, ub, vb, upper)
    for case in (2, 4):
        tc = _stress(case, ub, vb, upper)
        for a, b in zip(tc, t1):
            assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max(), case


@pytest.mark.parametrize("upper", [False, True])
def test_moeng_is_case4_linearised(upper):
    rng = np.random.default_rng(4)
    du = rng.standard_normal((24, 24))
    dv = rng.standard_normal((24, 24))
    d = []
    for eps in (1e-2, 1e-3):
        ub = 1.7 + eps * du
        vb = -0.4 + eps * dv
        t2 = _stress(2, ub, vb, upper)
        t4 = _stress(4, ub, vb, upper)
        d.append(max(np.abs(a - b).max() / np.abs(b).max() for a, b in zip(t2, t4)))
    assert d[0] < 5e-3, d
    assert d[1] < 2e-2 * d[0], d
