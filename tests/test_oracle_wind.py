"""CPU tests pinning the line-driven wind restatement (oracle/src/wind.c).

The golden fixtures were written by the compiled, unmodified reference
(problems wind_blondin / wind_plain, and wind_cvidl with the production
problem's UserDefBoundary from Test_Problems/LineDrivenWind/cv_idl/init.c;
tests/golden/make_golden.py) from the
grid and flux arrays stored beside them; the oracle must reproduce V, Uc,
the dt sequence, t and the next dt bit for bit.  Each fixture also states
which branches of the cooling and time-step logic it reaches, so a fixture
that silently stops covering (say) the stiff Cash-Karp path fails here.
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


def run_fixture(name):
    cfg, g = golden_io.load(name)
    grid = ({"x": g["x1"], "xl": g["x1l"], "xr": g["x1r"], "dx": g["dx1"]},
            {"x": g["x2"], "xl": g["x2l"], "xr": g["x2r"], "dx": g["dx2"]})
    sir, t0 = golden_io.sirocco(g)
    r = refrun.run_wind_oracle(cfg, g["V0"], int(g["nsteps"]), t0=t0,
                               fluxes=(g["flux_r"], g["flux_t"], g["flux_p"]), grid=grid,
                               sirocco=sir)
    return cfg, g, r, refrun.wind_counters()


@pytest.mark.parametrize("name", golden_io.names("wind"))
def test_wind_oracle_matches_reference_bitwise(name):
    cfg, g, r, _ = run_fixture(name)
    assert np.array_equal(r["V"], g["V"]), np.abs(r["V"] - g["V"]).max()
    assert np.array_equal(r["U"], g["U"]), np.abs(r["U"] - g["U"]).max()
    assert np.array_equal(r["dt_seq"], g["dt_seq"])
    assert r["t"] == float(g["t"]) and r["dt"] == float(g["dt"])


# branches each fixture must reach (counter > 0)
COVERAGE = {
    "wind_blondin_default_24x12": ["bl_solve", "dt_maxvar"],
    "wind_plain_default_24x12": ["rkf12_ok", "dt_maxvar"],
    "wind_blondin_mixed_24x12": ["bl_solve", "bl_widen", "below_minT", "c2p_fail"],
    "wind_blondin_mixed_cfl_24x12": ["bl_widen", "dt_cfl", "c2p_fail"],
    "wind_plain_mixed_24x12": ["rkf12_ok", "ck45_calls", "ck45_reject", "dt_cool"],
    "wind_cvidl_mixed_32x16": ["floor_cells", "ck45_calls", "c2p_fail"],
    "wind_cvidl_flat_mixed_32x16": ["floor_cells", "shock_zones", "ck45_calls"],
    "wind_cvidl_char_mixed_32x16": ["floor_cells", "shock_zones", "ck45_calls"],
    "wind_cvidl_prod_mixed_32x16": ["floor_cells", "shock_zones", "ck45_calls"],
    "wind_cvidl_prod_hll_mixed_32x16": ["floor_cells", "shock_zones", "ck45_calls"],
    "wind_plain_roe_mixed_32x16": ["rkf12_ok", "ck45_calls"],
    "wind_cvidl_prod_roe_mixed_32x16": ["floor_cells", "shock_zones", "ck45_calls"],
    "wind_plain_restart_32x16": ["rkf12_ok", "ck45_calls", "dt_cool"],
    "wind_blondin_restart_32x16": ["bl_solve", "below_minT"],
    "wind_plain_tabulated_32x16": ["rkf12_ok", "ck45_calls"],
    "wind_cvidl_prod_restart_32x16": ["floor_cells", "shock_zones", "ck45_calls"],
    "wind_cvidl_flat_midplane_32x64": ["floor_cells", "shock_zones", "midplane_cells"],
    "wind_cvidl_prod_midplane_32x64": ["floor_cells", "shock_zones", "midplane_cells"],
}


@pytest.mark.parametrize("name", sorted(COVERAGE))
def test_wind_fixture_branch_coverage(name):
    _, _, _, cnt = run_fixture(name)
    missing = [k for k in COVERAGE[name] if cnt[k] == 0]
    assert not missing, (missing, cnt)


def test_cvidl_fixture_pins_floors_and_sides():
    """The internal-boundary floors and the X1-beg userdef side change the
    result of the cv_idl fixture: dropping either one (in the oracle) breaks
    the bitwise match, so the fixture pins both.  The X2-beg hybrid side
    differs from a reflective one only in rho and p of the second ghost row,
    which no result can show with PLM + van Leer: the first ghost row equals
    the first interior row under both rules, so that row's slope difference
    is zero and the limiter returns a zero slope either way."""
    import dataclasses
    cfg, g = golden_io.load("wind_cvidl_mixed_32x16")
    grid = ({"x": g["x1"], "xl": g["x1l"], "xr": g["x1r"], "dx": g["dx1"]},
            {"x": g["x2"], "xl": g["x2l"], "xr": g["x2r"], "dx": g["dx2"]})
    fl = (g["flux_r"], g["flux_t"], g["flux_p"])
    for variant in (dataclasses.replace(cfg, dfloor=0.0),
                    dataclasses.replace(cfg, bc=(setups.BC_OUTFLOW,) + cfg.bc[1:])):
        r = refrun.run_wind_oracle(variant, g["V0"], int(g["nsteps"]), fluxes=fl, grid=grid)
        assert not np.array_equal(r["V"], g["V"])
    pole = dataclasses.replace(cfg, bc=cfg.bc[:2] + (setups.BC_REFLECTIVE,) + cfg.bc[3:])
    r = refrun.run_wind_oracle(pole, g["V0"], int(g["nsteps"]), fluxes=fl, grid=grid)
    assert np.array_equal(r["V"], g["V"])


def test_shock_flattening_fixture_pins_the_switches():
    """MULTID shock flattening changes the fixture's result, so the flags,
    the minmod slopes and the HLL fluxes are pinned by it."""
    import dataclasses
    cfg, g = golden_io.load("wind_cvidl_flat_mixed_32x16")
    grid = ({"x": g["x1"], "xl": g["x1l"], "xr": g["x1r"], "dx": g["dx1"]},
            {"x": g["x2"], "xl": g["x2l"], "xr": g["x2r"], "dx": g["dx2"]})
    fl = (g["flux_r"], g["flux_t"], g["flux_p"])
    r = refrun.run_wind_oracle(dataclasses.replace(cfg, shock_flattening=False), g["V0"],
                               int(g["nsteps"]), fluxes=fl, grid=grid)
    assert not np.array_equal(r["V"], g["V"])


def test_char_limiting_fixture_pins_the_projection():
    """Characteristic limiting changes the fixture's result, so the
    eigenvector projection and the primitive re-limiting are pinned by
    the bitwise match with the reference build."""
    import dataclasses
    cfg, g = golden_io.load("wind_cvidl_char_mixed_32x16")
    grid = ({"x": g["x1"], "xl": g["x1l"], "xr": g["x1r"], "dx": g["dx1"]},
            {"x": g["x2"], "xl": g["x2l"], "xr": g["x2r"], "dx": g["dx2"]})
    fl = (g["flux_r"], g["flux_t"], g["flux_p"])
    r = refrun.run_wind_oracle(dataclasses.replace(cfg, char_limiting=False), g["V0"],
                               int(g["nsteps"]), fluxes=fl, grid=grid)
    assert not np.array_equal(r["V"], g["V"])


def test_entropy_switch_fixture_pins_the_entropy_path():
    """With ENTROPY_SWITCH ALWAYS the pressure comes from the advected
    entropy; recovering it from the energy instead changes the fixture's
    result."""
    import dataclasses
    cfg, g = golden_io.load("wind_cvidl_prod_mixed_32x16")
    grid = ({"x": g["x1"], "xl": g["x1l"], "xr": g["x1r"], "dx": g["dx1"]},
            {"x": g["x2"], "xl": g["x2l"], "xr": g["x2r"], "dx": g["dx2"]})
    fl = (g["flux_r"], g["flux_t"], g["flux_p"])
    r = refrun.run_wind_oracle(dataclasses.replace(cfg, entropy_switch=False), g["V0"],
                               int(g["nsteps"]), fluxes=fl, grid=grid)
    assert not np.array_equal(r["V"], g["V"])


def test_hll_fixture_pins_the_solver():
    """The production problem's HLL solver: HLLC instead changes the
    fixture's result."""
    import dataclasses
    cfg, g = golden_io.load("wind_cvidl_prod_hll_mixed_32x16")
    grid = ({"x": g["x1"], "xl": g["x1l"], "xr": g["x1r"], "dx": g["dx1"]},
            {"x": g["x2"], "xl": g["x2l"], "xr": g["x2r"], "dx": g["dx2"]})
    fl = (g["flux_r"], g["flux_t"], g["flux_p"])
    r = refrun.run_wind_oracle(dataclasses.replace(cfg, solver=setups.HLLC), g["V0"],
                               int(g["nsteps"]), fluxes=fl, grid=grid)
    assert not np.array_equal(r["V"], g["V"])


def _restart_variant(name, cfg_change=None, **sir_change):
    """The oracle on a restart fixture with its config or SIROCCO inputs
    changed; True if the bitwise match with the reference survives."""
    cfg, g = golden_io.load(name)
    if cfg_change:
        cfg = dataclasses.replace(cfg, **cfg_change)
    sir, t0 = golden_io.sirocco(g)
    sir = dataclasses.replace(sir, **sir_change)
    grid = ({"x": g["x1"], "xl": g["x1l"], "xr": g["x1r"], "dx": g["dx1"]},
            {"x": g["x2"], "xl": g["x2l"], "xr": g["x2r"], "dx": g["dx2"]})
    r = refrun.run_wind_oracle(cfg, g["V0"], int(g["nsteps"]), t0=t0, sirocco=sir,
                               fluxes=(g["flux_r"], g["flux_t"], g["flux_p"]), grid=grid)
    return np.array_equal(r["V"], g["V"])


@pytest.mark.parametrize("name", ["wind_plain_restart_32x16", "wind_blondin_restart_32x16"])
def test_restart_fixtures_pin_prefactors_and_heatcool(name):
    """Each prefactor, the files' xi / T_r and the reader's floors change
    the restart fixtures: dropping the prefactors, swapping two of their
    columns (the two cooling terms brem_c / line_c, the Compton pair
    comp_h / comp_c), dropping xi / T_r, or handing the cooling the
    unfloored values all break the bitwise match with the reference."""
    cfg, g = golden_io.load(name)
    sir, _ = golden_io.sirocco(g)
    assert _restart_variant(name)
    assert not _restart_variant(name, pre=None)
    for a, b in ((3, 4), (0, 1), (0, 2)):
        pre = sir.pre.copy()
        pre[[a, b]] = pre[[b, a]]
        assert not _restart_variant(name, pre=pre), (a, b)
    assert not _restart_variant(name, xi=None, t_r=None)
    # the reader's floors reach cells of these fixtures
    assert (sir.xi < 1.0).any() and (sir.t_r < 1.e3).any()

    class NoFloor(setups.SiroccoInputs):
        def heatcool_floored(self):
            return self.xi, self.t_r
    cfg2, g2 = golden_io.load(name)
    grid = ({"x": g2["x1"], "xl": g2["x1l"], "xr": g2["x1r"], "dx": g2["dx1"]},
            {"x": g2["x2"], "xl": g2["x2l"], "xr": g2["x2r"], "dx": g2["dx2"]})
    r = refrun.run_wind_oracle(cfg2, g2["V0"], int(g2["nsteps"]), t0=float(g2["t0"]),
                               sirocco=NoFloor(xi=sir.xi, t_r=sir.t_r, pre=sir.pre),
                               fluxes=(g2["flux_r"], g2["flux_t"], g2["flux_p"]), grid=grid)
    assert not np.array_equal(r["V"], g2["V"])


def test_tabulated_fixture_pins_linterp():
    """The tabulated force multiplier fixture pins linterp(): the k, alpha
    form, the fits of another cell, the first or the last fit point (the
    values used below and above the fit range) and the interior points all
    change the result."""
    name = "wind_plain_tabulated_32x16"
    cfg, g = golden_io.load(name)
    sir, _ = golden_io.sirocco(g)
    assert _restart_variant(name)
    assert not _restart_variant(name, cfg_change={"krad": 0.2, "alpharad": -0.6})
    assert not _restart_variant(name, fm_m=np.roll(sir.fm_m, 1, axis=2))
    for q in (0, len(sir.fm_t) // 2, len(sir.fm_t) - 1):
        m = sir.fm_m.copy()
        m[q] *= 1.5
        assert not _restart_variant(name, fm_m=m), q


def test_linterp_closed_forms():
    """linterp() on a power law: with M = k t^a at the fit points the
    log-log interpolation is exact, so the tabulated force multiplier of a
    cell equals the k, alpha one wherever t lies inside the fit range; the
    tabulated run with fits of exactly k t^alpha over a range covering
    every t therefore agrees with the k, alpha run to rounding."""
    name = "wind_plain_tabulated_32x16"
    cfg, g = golden_io.load(name)
    sir, _ = golden_io.sirocco(g)
    t = np.logspace(-12.0, 8.0, 41)
    m = 0.2 * t[:, None, None] ** -0.6 * np.ones((1,) + sir.fm_m.shape[1:])
    grid = ({"x": g["x1"], "xl": g["x1l"], "xr": g["x1r"], "dx": g["dx1"]},
            {"x": g["x2"], "xl": g["x2l"], "xr": g["x2r"], "dx": g["dx2"]})
    fl = (g["flux_r"], g["flux_t"], g["flux_p"])
    a = refrun.run_wind_oracle(cfg, g["V0"], 4, fluxes=fl, grid=grid,
                               sirocco=setups.SiroccoInputs(fm_t=t, fm_m=m))
    b = refrun.run_wind_oracle(dataclasses.replace(cfg, krad=0.2, alpharad=-0.6), g["V0"], 4,
                               fluxes=fl, grid=grid)
    for nv in range(6):
        scale = np.abs(b["V"][nv]).max()
        np.testing.assert_allclose(a["V"][nv], b["V"][nv], rtol=1e-9, atol=1e-9 * scale)


def test_wind_fixture_inputs_match_generators():
    """The stored grid equals the generator's (so GPU tests that build the
    grid on their host feed the same geometry)."""
    cfg, g = golden_io.load("wind_plain_default_24x12")
    gr, gt = setups.wind_grid(cfg)
    np.testing.assert_allclose(g["x1"], gr["x"], rtol=1e-15)
    np.testing.assert_allclose(g["x2"], gt["x"], rtol=1e-15)


def test_log_grid_closed_form():
    """'l+' grid: faces r_k = r0 * (r1/r0)^(k/N) for r0 > 0, so consecutive
    width ratios are constant and the last face lands on r1."""
    g = setups.make_grid("l+", 1.0, 100.0, 64, 2)
    xl, xr = g["xl"][2:-2], g["xr"][2:-2]
    expect = 100.0 ** (np.arange(65) / 64.0)
    np.testing.assert_allclose(np.r_[xl, xr[-1]], expect, rtol=1e-12)
    ratio = g["dx"][3:-2] / g["dx"][2:-3]
    np.testing.assert_allclose(ratio, ratio[0], rtol=1e-12)


def test_wind_mass_budget_without_sources():
    """Mass is conserved up to the boundary fluxes: in a short run from a
    state at rest (zero radial/polar velocity, reflective polar walls) the
    total mass sum(rho dV) changes only through the two radial boundaries,
    and the change is far below one cell's mass."""
    cfg = setups.WindConfig(nx=(24, 12), cooling=setups.COOL_BLONDIN, first_dt=1e-5)
    V0 = setups.wind_ic(cfg, seed=3)
    r = refrun.run_wind_oracle(cfg, V0, 2)
    gr, gt = setups.wind_grid(cfg)
    rl, rr = gr["xl"][2:-2], gr["xr"][2:-2]
    tl, tr = gt["xl"][2:-2], gt["xr"][2:-2]
    dV = ((rr ** 3 - rl ** 3) / 3.0)[None, :] * np.abs(np.cos(tl) - np.cos(tr))[:, None]
    m0 = (V0[0, 0] * dV).sum()
    m1 = (r["V"][0, 0] * dV).sum()
    assert abs(m1 - m0) / m0 < 1e-3
    # the grid volume is the closed form (r1^3 - r0^3)/3 * (1 - cos(pi/2))
    assert math.isclose(dV.sum(), (100.0 ** 3 - 1.0) / 3.0 * (1.0 - math.cos(tr[-1])),
                        rel_tol=1e-12)


def test_roe_fixtures_pin_the_solver():
    """The Roe fixtures differ from the same runs with HLLC and with HLL:
    the solver choice is visible."""
    import dataclasses
    for name in ("wind_plain_roe_mixed_32x16", "wind_cvidl_prod_roe_mixed_32x16"):
        cfg, g = golden_io.load(name)
        grid = ({"x": g["x1"], "xl": g["x1l"], "xr": g["x1r"], "dx": g["dx1"]},
                {"x": g["x2"], "xl": g["x2l"], "xr": g["x2r"], "dx": g["dx2"]})
        fl = (g["flux_r"], g["flux_t"], g["flux_p"])
        for other in (setups.HLLC, setups.HLL):
            r = refrun.run_wind_oracle(dataclasses.replace(cfg, solver=other), g["V0"],
                                       int(g["nsteps"]), fluxes=fl, grid=grid)
            assert not np.array_equal(r["V"], g["V"]), (name, other)



@pytest.mark.parametrize("name", ["wind_cvidl_flat_midplane_32x64",
                                  "wind_cvidl_prod_midplane_32x64"])
def test_midplane_fixtures_pin_the_midplane_row(name):
    """At 64 polar zones the cv_idl side-0 pass resets the last interior
    row (MULTID runs with 3 ghost zones): each of the disk's parameters
    changes the fixture's result, so the reset is pinned; at 16 polar zones
    the row's centre and centroid differ by 2.6e-5 relative, above the
    pass's 1e-6 test, and the same parameters change nothing."""
    import dataclasses
    cfg, g = golden_io.load(name)
    grid = ({"x": g["x1"], "xl": g["x1l"], "xr": g["x1r"], "dx": g["dx1"]},
            {"x": g["x2"], "xl": g["x2l"], "xr": g["x2r"], "dx": g["dx2"]})
    fl = (g["flux_r"], g["flux_t"], g["flux_p"])
    for variant in (dataclasses.replace(cfg, rho_0=2.e-9), dataclasses.replace(cfg, rho_alpha=1.5),
                    dataclasses.replace(cfg, disk_mdot=1.e18)):
        r = refrun.run_wind_oracle(variant, g["V0"], int(g["nsteps"]), fluxes=fl, grid=grid)
        assert not np.array_equal(r["V"], g["V"])
    cfg16, g16 = golden_io.load("wind_cvidl_flat_mixed_32x16")
    grid16 = ({"x": g16["x1"], "xl": g16["x1l"], "xr": g16["x1r"], "dx": g16["dx1"]},
              {"x": g16["x2"], "xl": g16["x2l"], "xr": g16["x2r"], "dx": g16["dx2"]})
    r = refrun.run_wind_oracle(dataclasses.replace(cfg16, rho_0=2.e-9), g16["V0"],
                               int(g16["nsteps"]), grid=grid16,
                               fluxes=(g16["flux_r"], g16["flux_t"], g16["flux_p"]))
    assert np.array_equal(r["V"], g16["V"]) and refrun.wind_counters()["midplane_cells"] == 0
