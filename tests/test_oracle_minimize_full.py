"""Pins of the full-model minimize oracle (oracle/minimize_ref.py,
MinimizeFullOracle; CPU only).

The iteration composes pieces pinned elsewhere: the phi right-hand side
(tests/test_oracle_variational.py), the composition update
(tests/test_oracle_pins.py), the multiplier step and volume projection
(tests/test_oracle_minimize.py).  These tests pin what the composition adds:
a uniform matrix at the baseline composition is an exact stationary point
with zero chemical excess; the minimize Y cap keeps x_B at or below
minimize_xB_max_safe; and the pseudo-time flow lowers the total excess free
energy F_chem_excess + F_surf + F_el it reports."""
import math

import numpy as np

from cuda_sto_pf_amd.configs import base_params, config
from oracle import thermo as th
from oracle.minimize_ref import MinimizeFullOracle, MinimizeOptions
from oracle.step_ref import Params


def test_uniform_matrix_is_stationary_with_zero_excess():
    shape = (8, 8, 8)
    pd = base_params(shape, hetero=True, zero_mode=False, elastic=False)
    x0 = 0.021
    m = MinimizeFullOracle(Params(**pd), np.zeros(shape), np.full(shape, x0),
                           MinimizeOptions(dt=0.05, V0=0.0, xB_ref=0.0))
    assert m.x_ref == x0
    for _ in range(3):
        st = m.iterate()
        assert st["rms_dphi"] == 0.0 and st["rms_dY"] == 0.0
        assert abs(st["F_chem_excess"]) < 1e-15
        assert np.array_equal(m.o.phi, np.zeros(shape))
        assert np.abs(m.o.xB - x0).max() < 1e-15


def test_composition_stays_below_the_minimize_cap():
    pd, phi, xB, meta = config(2, 16)
    pd = dict(pd, zero_mode=0)
    cap = 0.0205   # just above the matrix composition: the flow presses against it
    opts = MinimizeOptions(dt=0.05, V0=0.0, xB_max_safe=cap)
    m = MinimizeFullOracle(Params(**pd), phi, xB, opts)
    assert m.o.xB.max() <= cap
    ycap = math.log(cap / (1.0 - cap))
    assert abs(m.y_cap - ycap) < 1e-15
    touched = 0
    for _ in range(4):
        m.iterate()
        assert m.o.Y.max() <= ycap and m.o.xB.max() <= cap * (1 + 1e-14)
        touched += int(np.count_nonzero(m.o.Y == ycap))
    assert touched > 0   # the cap was active


def test_pseudo_time_flow_lowers_the_total_excess_free_energy():
    pd, phi, xB, meta = config(2, 16)
    pd = dict(pd, zero_mode=0)
    m = MinimizeFullOracle(Params(**pd), phi, xB, MinimizeOptions(dt=0.05, V0=0.0))
    F = []
    for _ in range(8):
        F.append(m.iterate()["F_total_excess"])
    assert all(b < a for a, b in zip(F, F[1:])), F
    # the baseline is the far-field composition when given (ic_23d_xB_out > 0)
    m2 = MinimizeFullOracle(Params(**pd), phi, xB, MinimizeOptions(dt=0.05, xB_ref=0.03))
    g0 = (0.97 * float(th.mu_A_dimless(0.03, m2.o.T, pd["mu_reference_scale"]))
          + 0.03 * float(th.mu_B_dimless(0.03, m2.o.T, pd["mu_reference_scale"])))
    assert m2.x_ref == 0.03 and abs(m2.g_bulk0 - g0) < 1e-15
