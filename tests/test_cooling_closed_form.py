"""Closed-form checks of both cooling paths with bremsstrahlung only.

With every prefactor 0 except brem_c = 1 (prefactors.dat), the cooling
rate of a fully ionised cell is Lambda = -n_H n_e 3.3e-27 sqrt(T), and
T is proportional to the thermal energy E, so dE/dt = -B sqrt(E):

* the sub-cycled integrator (CoolingSource + Radiat, RKF12 / Cash-Karp)
  must follow sqrt(E(t)) = sqrt(E0) - B t / 2 to its error tolerance
  (n_e = 1.21 n_H there, cv_idl/radiat.c:97-104; the cut-off factor
  1 - 1/cosh((T/1e4)^12) is exactly 1 at T >= 1e5 K);
* BLONDIN's implicit trapezoid step E_f - E - dt (Lambda(T) + Lambda(T_f))/2
  = 0 (Src/Cooling/BLONDIN/cooling.c:233-237) becomes a quadratic in
  sqrt(T_f) (n_e/n_H = 1.21 above 3.3e4 K), whose root Brent must find to
  its 1 K tolerance.

Neither check uses the reference: they pin the prefactor wiring (a wrong
column would leave another rate in, or drop this one) and the integrators
by mathematics.  The oracle is checked here; tests/test_gpu_wind.py holds
the same checks on the device.
"""
import dataclasses

import numpy as np
import pytest

import setups
from oracle import refrun

KB, MP = setups.CONST["kB"], setups.CONST["mp"]


def brem_case(cooling):
    cfg = setups.WindConfig(nx=(8, 4), cooling=cooling)
    nt, nr = cfg.nx[1], cfg.nx[0]
    V = np.zeros(cfg.shape)
    kelvin = cfg.unit_velocity ** 2 * setups.CONST["amu"] / KB
    rho = np.linspace(1e-13, 4e-13, nt * nr).reshape(nt, nr)     # g/cm^3
    T0 = np.linspace(2e6, 2e7, nt * nr).reshape(nt, nr)[::-1]    # K
    V[setups.RHO, 0] = rho / cfg.unit_density
    V[setups.PRS, 0] = V[setups.RHO, 0] * T0 / (kelvin * cfg.mu)
    V[setups.TRC] = 1.0
    pre = np.zeros((6, nt, nr))
    pre[3] = 1.0                        # brem_c (prefactors.dat column order)
    return cfg, V, pre, kelvin


def analytic_plain(cfg, V, dt, kelvin):
    """p after dt under dE/dt = -B sqrt(E), code units."""
    g = cfg.gamma
    rho = V[setups.RHO, 0]
    E0 = V[setups.PRS, 0] / (g - 1.0)
    nH = rho * cfg.unit_density / (1.43 * MP)
    E_cost = cfg.unit_length / cfg.unit_density / cfg.unit_velocity ** 3
    a = (g - 1.0) * kelvin * cfg.mu / rho            # T = a E
    B = 1.21 * nH * nH * 3.3e-27 * E_cost * np.sqrt(a)
    return (np.sqrt(E0) - 0.5 * B * dt) ** 2 * (g - 1.0)


def analytic_blondin(cfg, V, dt):
    """p after one BLONDIN step (the trapezoid root), code units."""
    g = cfg.gamma
    UP = cfg.unit_density * cfg.unit_velocity ** 2
    UT = cfg.unit_length / cfg.unit_velocity
    rho = V[setups.RHO, 0] * cfg.unit_density
    E = V[setups.PRS, 0] * UP / (g - 1.0)
    n = rho / (cfg.mu * MP)
    nH = rho / (1.43 * MP)
    c1 = n * KB / (2.0 / 3.0)                        # E = c1 T
    D = nH * nH * 1.21 * 3.3e-27                     # -Lambda / sqrt(T)
    T = E / c1
    h = 0.5 * dt * UT
    # c1 x^2 + h D x - (E - h D sqrt(T)) = 0, x = sqrt(T_f)
    cc = E - h * D * np.sqrt(T)
    x = (-h * D + np.sqrt((h * D) ** 2 + 4.0 * c1 * cc)) / (2.0 * c1)
    return (c1 * x * x) * (g - 1.0) / UP, x * x


@pytest.fixture(scope="module", autouse=True)
def _build():
    refrun.build(with_reference=False)


@pytest.mark.parametrize("dt", [1e-2, 1.0, 30.0])
def test_plain_integrator_follows_sqrt_law(dt):
    cfg, V, pre, kelvin = brem_case(setups.COOL_PLAIN)
    r = refrun.run_wind_cool_oracle(cfg, V, dt, t=10.0, pre=pre)
    p_an = analytic_plain(cfg, V, dt, kelvin)
    assert (p_an > 0.05 * V[setups.PRS, 0]).all()      # stays hot: the law holds throughout
    np.testing.assert_allclose(r["V"][setups.PRS, 0], p_an, rtol=1e-4)
    if dt == 30.0:
        assert refrun.wind_counters()["ck45_calls"] > 0     # the stiff path is covered


def test_plain_integrator_wrong_prefactor_column_fails():
    """The check is sharp: the same rates under line_c instead of brem_c
    (a transposed column) miss the sqrt law by far."""
    cfg, V, pre, kelvin = brem_case(setups.COOL_PLAIN)
    pre = pre[[0, 1, 2, 4, 3, 5]]
    r = refrun.run_wind_cool_oracle(cfg, V, 30.0, t=10.0, pre=pre)
    p0, p_an = V[setups.PRS, 0], analytic_plain(cfg, V, 30.0, kelvin)
    drop_err = np.abs((p0 - r["V"][setups.PRS, 0]) / (p0 - p_an) - 1.0)
    assert drop_err.min() > 0.5


@pytest.mark.parametrize("dt", [1e-2, 1.0, 30.0])
def test_blondin_step_is_the_trapezoid_root(dt):
    cfg, V, pre, _ = brem_case(setups.COOL_BLONDIN)
    r = refrun.run_wind_cool_oracle(cfg, V, dt, t=10.0, pre=pre)
    p_an, Tf = analytic_blondin(cfg, V, dt)
    assert (Tf > 3.3e4).all()
    # Brent stops within 1 K of the root
    dT = np.abs(r["V"][setups.PRS, 0] / p_an - 1.0) * Tf
    assert dT.max() <= 1.0, dT.max()


def test_exact_cooling_time_closed_form():
    """The oracle's exact cooling time (the integral of dE / f(E) that judges
    the surrogate) against the bremsstrahlung law: with only brem_c left,
    dE/dt = -B sqrt(E), so the time from E0 to E1 is 2 (sqrt(E0) - sqrt(E1))
    / B; a transposed prefactor (line_c for brem_c) misses it."""
    cfg, V, pre, kelvin = brem_case(setups.COOL_PLAIN)
    g = cfg.gamma
    E_cost = cfg.unit_length / cfg.unit_density / cfg.unit_velocity ** 3
    pre5 = pre[:5, 0, 0]
    for q in range(3):
        rho = float(V[setups.RHO, 0].ravel()[q])
        p0 = float(V[setups.PRS, 0].ravel()[q])
        nH = rho * cfg.unit_density / (1.43 * MP)
        a = (g - 1.0) * kelvin * cfg.mu / rho                 # T = a E
        B = 1.21 * nH * nH * 3.3e-27 * E_cost * np.sqrt(a)
        E0 = p0 / (g - 1.0)
        for frac in (0.9, 0.5, 0.1):
            p1 = frac * p0                                   # T stays above 1e5 K
            assert p1 / (g - 1.0) * a > 1.e5
            E1 = p1 / (g - 1.0)
            t_an = 2.0 * (np.sqrt(E0) - np.sqrt(E1)) / B
            t = refrun.wind_cool_time(cfg, rho, p0, 1.0, p1, 1e-12 * t_an, pre=pre5)
            assert abs(t / t_an - 1.0) < 1e-9, (q, frac, t, t_an)
        swapped = pre5.copy()
        swapped[[3, 4]] = swapped[[4, 3]]
        t_sw = refrun.wind_cool_time(cfg, rho, p0, 1.0, 0.5 * p0, 1e-9, pre=swapped)
        assert not abs(t_sw / (2.0 * (np.sqrt(E0) - np.sqrt(0.5 * E0)) / B) - 1.0) < 1e-3


def test_exact_cooling_time_is_additive():
    """With every rate on (unit prefactors), the time from p0 to p2 is the
    time to an intermediate p1 plus the time from there (the integral's own
    consistency, which a wrong sign, bound or quadrature rule breaks), and
    it is 0 from a state to itself."""
    cfg = setups.WindConfig(nx=(8, 4), cooling=setups.COOL_PLAIN)
    kelvin = cfg.unit_velocity ** 2 * setups.CONST["amu"] / KB
    for rho_cgs, T0, xi in ((1e-12, 3e6, 10.0), (3e-11, 1e7, 300.0), (1e-13, 8e5, 1e4)):
        rho = rho_cgs / cfg.unit_density
        p0 = rho * T0 / (kelvin * cfg.mu)
        t_all = refrun.wind_cool_time(cfg, rho, p0, xi, 0.3 * p0, 1e-14)
        if not np.isfinite(t_all):
            continue
        t_a = refrun.wind_cool_time(cfg, rho, p0, xi, 0.6 * p0, 1e-14)
        t_b = refrun.wind_cool_time(cfg, rho, 0.6 * p0, xi, 0.3 * p0, 1e-14)
        assert t_all > 0.0 and abs((t_a + t_b) / t_all - 1.0) < 1e-9, (rho_cgs, t_a, t_b, t_all)
        assert refrun.wind_cool_time(cfg, rho, p0, xi, p0, 1e-14) == 0.0
