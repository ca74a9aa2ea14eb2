"""The oracle's largest-signal-speed rule (oracle/cfl.c), pinned by closed
forms: states at rest and moving, the MHD fast speed across B against the
speed the MHD Rusanov flux itself returns, the two-fluid maximum over its
signals, and NaN propagation."""
import math

import numpy as np
import pytest

from oracle import oracle as O

from mojoxm_amd import physics as ph


def test_euler_rest_and_moving():
    e = ph.Euler(gamma=1.4)
    q = np.array([1.2, 0.0, 0.0, 0.0, 2.0 / 0.4])
    assert O.max_speed(e.phys_id, e.params(), q) == pytest.approx(math.sqrt(1.4 * 2.0 / 1.2), rel=1e-15)
    u = np.array([0.3, -0.4, 0.0])
    q = np.array([1.2, *(1.2 * u), 2.0 / 0.4 + 0.5 * 1.2 * 0.25])
    assert O.max_speed(e.phys_id, e.params(), q) == pytest.approx(0.5 + math.sqrt(1.4 * 2.0 / 1.2), rel=1e-14)
    e2 = ph.Euler2D(gamma=1.4)
    q2 = np.array([1.2, 1.2 * 0.3, -1.2 * 0.4, 2.0 / 0.4 + 0.5 * 1.2 * 0.25])
    assert O.max_speed(e2.phys_id, e2.params(), q2) == pytest.approx(0.5 + math.sqrt(1.4 * 2.0 / 1.2), rel=1e-14)


def test_density_and_pressure_floors():
    e = ph.Euler(gamma=1.4, min_density=0.1, min_pressure=0.05)
    q = np.array([0.01, 0.0, 0.0, 0.0, -1.0])  # both floored
    assert O.max_speed(e.phys_id, e.params(), q) == pytest.approx(math.sqrt(1.4 * 0.05 / 0.1), rel=1e-15)


@pytest.mark.parametrize("phys", [ph.ShallowWater2D(g=9.81), ph.ShallowWater(g=9.81)])
def test_shallow_water(phys):
    q = np.array([2.0, 2.0 * 0.6, 2.0 * 0.8])
    assert O.max_speed(phys.phys_id, phys.params(), q) == pytest.approx(1.0 + math.sqrt(9.81 * 2.0), rel=1e-15)


def test_mhd_fast_speed_matches_the_flux_speed():
    """With u along n and B across n the Rusanov flux's speed |u.n| + c_f(n)
    is the largest signal speed over all directions."""
    g = 5.0 / 3.0
    mhd = ph.IdealMHD(gamma=g, c_h=0.0)
    rng = np.random.default_rng(1)
    for _ in range(10):
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        B = np.cross(n, rng.standard_normal(3))
        rho, p, un = rng.uniform(0.8, 1.2), rng.uniform(0.5, 1.5), rng.uniform(-0.5, 0.5)
        u = un * n
        q = np.array([rho, *(rho * u), p / (g - 1) + 0.5 * rho * u @ u + 0.5 * B @ B, *B, 0.0])
        _, speed = O.numerical_flux(mhd.phys_id, mhd.params(), q, q, n)
        assert O.max_speed(mhd.phys_id, mhd.params(), q) == pytest.approx(speed, rel=1e-13)
        assert speed == pytest.approx(abs(un) + math.sqrt((g * p + B @ B) / rho), rel=1e-13)
    # c_h dominates a slow state
    mhd = ph.IdealMHD(gamma=g, c_h=7.0)
    q = np.array([1.0, 0, 0, 0, 1.5, 0.1, 0, 0, 0])
    assert O.max_speed(mhd.phys_id, mhd.params(), q) == 7.0
    m2 = ph.MHDGLM2D(gamma=g, c_h=0.5)
    q2 = np.array([1.0, 0.3, 0.4, 0.6, 0.8, 0.1 / (g - 1) + 0.5 * 0.25 + 0.5, 0.0])
    assert O.max_speed(m2.phys_id, m2.params(), q2) == pytest.approx(0.5 + math.sqrt((g * 0.1 + 1.0) / 1.0), rel=1e-13)


def test_maxwell_and_advection():
    for mw in (ph.Maxwell(c=2.5), ph.Maxwell2D(c=2.5)):
        assert O.max_speed(mw.phys_id, mw.params(), np.ones(6)) == 2.5
    a = ph.Advection(1.0, -2.0, 2.0)
    assert O.max_speed(a.phys_id, a.params(), np.ones(1)) == pytest.approx(3.0, rel=1e-15)
    a2 = ph.Advection2D(0.6, -0.8)
    assert O.max_speed(a2.phys_id, a2.params(), np.ones(1)) == pytest.approx(1.0, rel=1e-15)


def test_two_fluid_largest_signal():
    q = np.zeros(17)
    q[0], q[1], q[4] = 1.0, 0.5, 1.0 / (5.0 / 3.0 - 1) + 0.125   # electrons: u 0.5, p 1
    q[5], q[9] = 4.0, 0.5 / (1.4 - 1)                              # ions at rest, p 0.5
    ce = 0.5 + math.sqrt(5.0 / 3.0)
    for cl, ch, want in ((10.0, 1.0, 10.0), (0.5, 3.0, 3.0), (0.1, 0.2, ce)):
        tf = ph.FiveMomentTwoFluid(gamma_e=5.0 / 3.0, gamma_i=1.4, c_light=cl, c_h=ch)
        assert O.max_speed(tf.phys_id, tf.params(), q) == pytest.approx(want, rel=1e-14)


def test_nan_propagates():
    e = ph.Euler()
    q = np.array([[[1.0, 0.0, 0.0, 0.0, 2.5], [np.nan, 0.0, 0.0, 0.0, 2.5]]])
    assert math.isnan(O.max_wave_speed(e.phys_id, e.params(), q))
    mhd = ph.IdealMHD(c_h=5.0)
    qm = np.array([np.nan, 0, 0, 0, 1, 0, 0, 0, 0])
    assert math.isnan(O.max_speed(mhd.phys_id, mhd.params(), qm))
