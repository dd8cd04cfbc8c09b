"""A SIROCCO coupling cycle through the library's public calls, end to end
on the GPU: the cycle's files (directional fluxes, force-multiplier fits,
py_heatcool.dat, prefactors.dat) are read with the library's readers, the
production set-up runs from them at a restart time, its output (V and
the 23 user variables) is written as .dbl / dbl.out / restart.out, and a
second context restarted from that output continues the run.  Checked
against the oracle given the same inputs (1e-12, as the wind step), the
restart against the uninterrupted run (bit for bit)."""
import dataclasses

import numpy as np
import pytest

import setups
from oracle import refrun

pytestmark = pytest.mark.gpu

RTOL = 1e-12


@pytest.fixture(scope="module")
def pmx():
    import plutomx
    plutomx.load_library()
    refrun.build(with_reference=False)
    return plutomx


def close(a, b, what):
    for nv in range(b.shape[0]):
        scale = np.max(np.abs(b[nv]))
        bad = np.abs(a[nv] - b[nv]) > RTOL * np.abs(b[nv]) + RTOL * scale
        assert not bad.any(), (what, nv, int(bad.sum()))


def test_sirocco_cycle_end_to_end(pmx, tmp_path):
    cfg = dataclasses.replace(setups.cvidl_wind(48, 24, dfloor=1.e-15), first_dt=1.e-3,
                              shock_flattening=True, char_limiting=True, entropy_switch=True,
                              solver=setups.HLL, krad=999.0, alpharad=999.0)
    V0 = setups.wind_ic_mixed(cfg, rho0_cgs=1.e-11, t_range=(1.e2, 3.e7))
    inp = setups.sirocco_inputs(cfg, V0, fm_points=12, fm_range=(1e-4, 1e-2))
    cyc = tmp_path / "cycle"
    cyc.mkdir()
    setups.write_flux_files(cfg, cyc)
    setups.write_sirocco_files(cfg, cyc, inp, V0)

    # the cycle's files, read as a restart does
    r = pmx.read_sirocco_files(cfg, cyc, V0)
    got = r["inputs"]
    n1, n2, t0 = 4, 4, 10.0
    out = tmp_path / "out"
    out.mkdir()
    names = list(pmx.PRIM_NAMES) + list(setups.USER_VARS)
    with pmx.Wind(cfg, fluxes=r["fluxes"], sirocco=got) as w:
        w.set_prim(V0)
        w.set_sirocco(got)
        w.set_time(t0, cfg.first_dt, 0)
        w.run(n1)
        w.snapshot(user_vars=True)
        s = w.snapshot_wait()
        data = np.concatenate([s["V"], s["uv"]])
        pmx.write_dbl(out, 0, data, names, s["t"], s["dt"], s["step"])
        pmx.write_restart(out, 0, s["step"], [0] + [-1] * 15, s["t"], s["dt"])
        w.run(n2)
        w.sync()
        V_end = w.get_prim()

    # the oracle from the same inputs: the state at the output and at the end
    orc1 = refrun.run_wind_oracle(cfg, V0, n1, t0=t0, fluxes=r["fluxes"], sirocco=got,
                                  user_vars=True)
    close(s["V"][:, 0], orc1["V"][:, 0], "V at the output")
    close(s["uv"][:, 0], orc1["uv"], "user variables")
    orc = refrun.run_wind_oracle(cfg, V0, n1 + n2, t0=t0, fluxes=r["fluxes"], sirocco=got)
    close(V_end[:, 0], orc["V"][:, 0], "V at the end")

    # restart from the written output: the same end state, bit for bit
    rs = pmx.read_restart(out, 0, 6, (1, cfg.nx[1], cfg.nx[0]))
    assert rs["nstep"] == n1 and rs["t"] == s["t"] and rs["dt"] == s["dt"]
    with pmx.Wind(cfg, fluxes=r["fluxes"], sirocco=got) as w2:
        w2.set_prim(rs["data"])
        w2.set_sirocco(got)
        w2.set_time(rs["t"], rs["dt"], rs["nstep"])
        w2.run(n2)
        w2.sync()
        assert np.array_equal(w2.get_prim(), V_end)
