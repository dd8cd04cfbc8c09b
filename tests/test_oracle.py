"""CPU tests pinning the oracle (the plain-C restatement of the hydro step).

* Golden fixtures written by the compiled, unmodified reference
  (tests/golden/make_golden.py): the oracle must reproduce V, Uc, the dt
  sequence, t and the next dt bit for bit.
* Closed forms and invariants that hold whatever the implementation:
  uniform states are stationary, periodic runs conserve mass, momentum and
  energy, supersonic interfaces take the upwind flux, and a Sod shock tube
  converges to Toro's exact solution.
"""
import dataclasses
import math

import numpy as np
import pytest

import golden_io
import setups
from oracle import refrun


@pytest.fixture(scope="module", autouse=True)
def _build():
    refrun.build(with_reference=False)


@pytest.mark.parametrize("name", golden_io.names())
def test_oracle_matches_reference_bitwise(name):
    cfg, g = golden_io.load(name)
    r = refrun.run_oracle(cfg, g["V0"], int(g["nsteps"]))
    assert np.array_equal(r["V"], g["V"]), np.abs(r["V"] - g["V"]).max()
    assert np.array_equal(r["U"], g["U"]), np.abs(r["U"] - g["U"]).max()
    assert np.array_equal(r["dt_seq"], g["dt_seq"])
    assert r["t"] == float(g["t"]) and r["dt"] == float(g["dt"])


CASES = [
    setups.HydroConfig(dims=2, nx=(16, 12, 1), recon=setups.PLM, limiter=setups.LIM_VANLEER),
    setups.HydroConfig(dims=3, nx=(8, 6, 7), recon=setups.PPM, rk=3, solver=setups.ROE),
    setups.HydroConfig(dims=3, nx=(6, 8, 5), recon=setups.PLM, limiter=setups.LIM_DEFAULT,
                       solver=setups.HLLC, rk=3),
]


@pytest.mark.parametrize("cfg", CASES)
def test_uniform_state_is_stationary(cfg):
    V0 = np.empty(cfg.shape)
    V0[0], V0[1], V0[2], V0[3], V0[4] = 1.3, 0.2, -0.1, 0.05, 0.7
    r = refrun.run_oracle(cfg, V0, 3)
    np.testing.assert_allclose(r["V"], V0, rtol=1e-14, atol=1e-15)


def V0_size(cfg):
    return cfg.ncells


@pytest.mark.parametrize("cfg", CASES)
def test_periodic_conservation(cfg):
    cfg = dataclasses.replace(cfg, bc=(setups.BC_PERIODIC,) * 6, first_dt=1e-3)
    # smooth positive state (no floor is ever applied, so the update is
    # exactly conservative up to round-off)
    ph = np.arange(V0_size(cfg)).reshape(cfg.shape[1:]) * 0.37
    V0 = np.empty(cfg.shape)
    V0[0] = 1.0 + 0.2 * np.sin(ph)
    V0[1], V0[2], V0[3] = 0.3 * np.cos(ph), 0.2 * np.sin(1.3 * ph), 0.1 * np.cos(0.7 * ph)
    V0[4] = 1.0 + 0.3 * np.cos(0.9 * ph)
    U0 = np.empty_like(V0)
    lib = refrun.lib()
    for idx in np.ndindex(V0[0].shape):
        v = np.ascontiguousarray(V0[(slice(None),) + idx])
        u = np.empty(5)
        lib.orc_prim_to_cons(cfg.gamma, refrun._ptr(v), refrun._ptr(u))
        U0[(slice(None),) + idx] = u
    r = refrun.run_oracle(cfg, V0, 5)
    for nv in range(5):
        s0, s1 = U0[nv].sum(), r["U"][nv].sum()
        scale = np.abs(U0[nv]).sum()
        assert abs(s1 - s0) <= 1e-13 * scale, (nv, s0, s1)


@pytest.mark.parametrize("solver", [setups.HLLC, setups.ROE])
@pytest.mark.parametrize("direction", [0, 1, 2])
def test_supersonic_interface_is_upwind(solver, direction):
    gamma = 5.0 / 3.0
    vL = np.array([1.0, 0.3, -0.2, 0.1, 0.5])
    vR = np.array([0.8, 0.25, 0.1, -0.3, 0.4])
    vL[1 + direction] = vR[1 + direction] = 5.0      # Mach ~5 to the right
    flux, p, cmax = np.empty(5), np.empty(1), np.empty(1)
    refrun.lib().orc_riemann(solver, gamma, direction, refrun._ptr(vL), refrun._ptr(vR),
                             refrun._ptr(flux), refrun._ptr(p), refrun._ptr(cmax))
    u = np.empty(5)
    refrun.lib().orc_prim_to_cons(gamma, refrun._ptr(vL), refrun._ptr(u))
    vn = vL[1 + direction]
    exact = np.array([u[1 + direction], u[1] * vn, u[2] * vn, u[3] * vn, (u[4] + vL[4]) * vn])
    if solver == setups.HLLC:
        assert np.array_equal(flux, exact) and p[0] == vL[4]
    else:  # Roe splits the pressure: compare the total normal-momentum flux
        total, ref = flux.copy(), exact.copy()
        total[1 + direction] += p[0]
        ref[1 + direction] += vL[4]
        np.testing.assert_allclose(total, ref, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("solver", [setups.HLLC, setups.ROE])
def test_equal_states_give_physical_flux(solver):
    gamma = 1.4
    v = np.array([1.2, -0.4, 0.3, 0.2, 0.9])
    flux, p, cmax = np.empty(5), np.empty(1), np.empty(1)
    refrun.lib().orc_riemann(solver, gamma, 0, refrun._ptr(v), refrun._ptr(v.copy()),
                             refrun._ptr(flux), refrun._ptr(p), refrun._ptr(cmax))
    u = np.empty(5)
    refrun.lib().orc_prim_to_cons(gamma, refrun._ptr(v), refrun._ptr(u))
    exact = np.array([u[1], u[1] * v[1], u[2] * v[1], u[3] * v[1], (u[4] + v[4]) * v[1]])
    np.testing.assert_allclose(flux, exact, rtol=1e-13, atol=1e-15)
    assert p[0] == v[4]
    a = math.sqrt(gamma * v[4] / v[0])
    np.testing.assert_allclose(cmax[0], abs(v[1]) + a, rtol=1e-14)


# ---------------------------------------------------------------------
# Sod shock tube against the exact solution (Toro, ch. 4)
# ---------------------------------------------------------------------
def _sod_exact(x, t, gamma=1.4, left=(1.0, 0.0, 1.0), right=(0.125, 0.0, 0.1)):
    rl, ul, pl = left
    rr, ur, pr = right
    cl, cr = math.sqrt(gamma * pl / rl), math.sqrt(gamma * pr / rr)

    def f(p, r, pk, ck):
        if p > pk:
            A, B = 2.0 / ((gamma + 1) * r), (gamma - 1) / (gamma + 1) * pk
            return (p - pk) * math.sqrt(A / (p + B))
        return 2 * ck / (gamma - 1) * ((p / pk) ** ((gamma - 1) / (2 * gamma)) - 1)

    lo, hi = 1e-8, 10.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid, rl, pl, cl) + f(mid, rr, pr, cr) + ur - ul > 0:
            hi = mid
        else:
            lo = mid
    ps = 0.5 * (lo + hi)
    us = 0.5 * (ul + ur) + 0.5 * (f(ps, rr, pr, cr) - f(ps, rl, pl, cl))
    rsl = rl * (ps / pl) ** (1 / gamma)
    csl = cl * (ps / pl) ** ((gamma - 1) / (2 * gamma))
    g1 = (gamma - 1) / (gamma + 1)
    rsr = rr * (ps / pr + g1) / (g1 * ps / pr + 1)
    sshock = ur + cr * math.sqrt((gamma + 1) / (2 * gamma) * ps / pr + (gamma - 1) / (2 * gamma))
    rho = np.empty_like(x)
    for n, xi in enumerate(x / t):
        if xi < ul - cl:
            rho[n] = rl
        elif xi < us - csl:
            c = 2 / (gamma + 1) * (cl + (gamma - 1) / 2 * (ul - xi))
            rho[n] = rl * (c / cl) ** (2 / (gamma - 1))
        elif xi < us:
            rho[n] = rsl
        elif xi < sshock:
            rho[n] = rsr
        else:
            rho[n] = rr
    return rho


@pytest.mark.parametrize("recon,solver", [(setups.PLM, setups.HLLC), (setups.PPM, setups.HLLC),
                                          (setups.PLM, setups.ROE)])
def test_sod_converges_to_exact_solution(recon, solver):
    n = 200
    cfg = setups.HydroConfig(dims=2, nx=(n, 4, 1), recon=recon, limiter=setups.LIM_VANLEER,
                             solver=solver, rk=2, gamma=1.4, cfl=0.4, first_dt=1e-4,
                             bc=(setups.BC_OUTFLOW, setups.BC_OUTFLOW, setups.BC_PERIODIC,
                                 setups.BC_PERIODIC, 0, 0),
                             xbeg=(0.0, 0.0, 0.0), xend=(1.0, 0.02, 1.0))
    x = setups.cell_centers(cfg, 0)
    V0 = np.zeros(cfg.shape)
    left = x < 0.5
    V0[0] = np.where(left, 1.0, 0.125)[None, None, :]
    V0[4] = np.where(left, 1.0, 0.1)[None, None, :]
    t, dt, V = 0.0, cfg.first_dt, V0
    while t < 0.2 - 1e-12:
        r = refrun.run_oracle(cfg, V, 1, dt0=min(dt, 0.2 - t), t0=t)
        V, t, dt = r["V"], r["t"], r["dt"]
    rho = V[0, 0, 1]
    exact = _sod_exact(x - 0.5, 0.2)
    l1 = np.abs(rho - exact).mean()
    assert l1 < 6e-3, l1
    # states far from the waves are untouched
    assert abs(rho[5] - 1.0) < 1e-12 and abs(rho[-5] - 0.125) < 1e-12
