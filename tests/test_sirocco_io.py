"""CPU tests of the SIROCCO file readers (libplutomx pmx_sirocco_*, host
code: no device is touched).

The readers restate read_sirocco_fluxes / read_sirocco_heatcool
(Src/LineDriven/line_connect.c:43-497).  They are pinned three ways:

* files written from known arrays with 17 significant digits read back to
  exactly those arrays (xi / T_r after the floors at 1 and 1e3 K);
* the reference's rules for what a record matches (relative tolerance 1e-6
  on r and theta for fluxes and fits, 1e-5 relative to the record's own r
  and theta for py_heatcool / prefactors), for missing files and for
  malformed records, on files built to straddle each rule;
* end to end against the compiled reference (skipped where oracle/_ref is
  not built): the reference reads a set of files with records moved off
  the grid inside and outside the tolerances, the oracle runs on what this
  library's readers make of the same files, and the two states agree bit
  for bit.
"""
import dataclasses
import numpy as np
import pytest

import setups
from oracle import refrun


@pytest.fixture(scope="module")
def pmx():
    import plutomx
    plutomx.load_library()
    return plutomx


CFG = setups.WindConfig(nx=(32, 16))


def state():
    return setups.wind_ic_mixed(CFG, rho0_cgs=1.e-11, t_range=(1.e2, 3.e7))


def write_all(d, V, inp):
    setups.write_flux_files(CFG, d)
    setups.write_sirocco_files(CFG, d, inp, V)


def test_readers_round_trip(pmx, tmp_path):
    V = state()
    inp = setups.sirocco_inputs(CFG, V, fm_points=12, fm_range=(1e-4, 1e-2))
    write_all(tmp_path, V, inp)
    r = pmx.read_sirocco_files(CFG, tmp_path, V)
    n = CFG.nx[0] * CFG.nx[1]
    assert r["found"] == {"fluxes": 7, "fm": 1, "heatcool": 1, "prefactors": 1}
    assert r["matched"] == {"fluxes": [n, n, n], "fm": n, "heatcool": n, "prefactors": n}
    for a, b in zip(r["fluxes"], setups.wind_fluxes(CFG)):
        assert np.array_equal(a, b)
    got = r["inputs"]
    xi, t_r = inp.heatcool_floored()
    assert np.array_equal(got.xi, xi) and np.array_equal(got.t_r, t_r)
    assert (inp.xi < 1.0).any() and (inp.t_r < 1.e3).any()   # the floors are exercised
    assert np.array_equal(got.pre, inp.pre)
    assert np.array_equal(got.fm_t, inp.fm_t) and np.array_equal(got.fm_m, inp.fm_m)
    assert r["negative_prefactors"] == 0


def test_missing_files_give_the_start_up_values(pmx, tmp_path):
    """Without py_heatcool.dat the restart read sets xi = L_x / (n_H r^2)
    and T_r = T_x (line_connect.c:313-326); without prefactors.dat every
    prefactor is 1 (:366-380); without flux or fit files nothing is read."""
    V = state()
    r = pmx.read_sirocco_files(CFG, tmp_path, V)
    assert r["fluxes"] is None and r["inputs"].fm_t is None
    assert r["found"] == {"heatcool": 0, "prefactors": 0}
    assert (r["inputs"].pre == 1.0).all()
    gr, _ = setups.wind_grid(CFG)
    rad = gr["x"][2:-2][None, :] * CFG.unit_length
    nH = V[setups.RHO, 0] * CFG.unit_density / (1.43 * setups.CONST["mp"])
    np.testing.assert_allclose(r["inputs"].xi, CFG.l_star * CFG.f_x / nH / rad ** 2, rtol=1e-15)
    assert (r["inputs"].t_r == CFG.t_x).all()


def _shift_records(path, frac, which, col_r):
    """Multiply the r column of the listed record lines (0-based, after the
    header lines) by (1 + frac)."""
    lines = path.read_text().splitlines(keepends=True)
    head = 2 if path.name.startswith("directional") else 1
    for q in which:
        w = lines[head + q].split()
        w[col_r] = f"{float(w[col_r]) * (1.0 + frac):.17e}"
        lines[head + q] = " ".join(w) + "\n"
    path.write_text("".join(lines))


def test_match_tolerances(pmx, tmp_path):
    """A record whose r is off by a relative 5e-6 still matches in
    py_heatcool / prefactors (tolerance 1e-5) but not in the flux or fit
    files (1e-6); off by 3e-5 it matches nowhere.  Unmatched records are
    skipped: their cells keep 1.0 (prefactors) or the caller's values."""
    V = state()
    inp = setups.sirocco_inputs(CFG, V, fm_points=4)
    write_all(tmp_path, V, inp)
    near, far = [3, 40], [5, 77]
    for name, col in (("prefactors.dat", 1), ("py_heatcool.dat", 2),
                      ("directional_flux_r.dat", 3), ("M_UV_data.dat", 2)):
        _shift_records(tmp_path / name, 5e-6, near, col)
        _shift_records(tmp_path / name, 3e-5, far, col)
    r = pmx.read_sirocco_files(CFG, tmp_path, V)
    n = CFG.nx[0] * CFG.nx[1]
    assert r["matched"]["prefactors"] == n - 2 and r["matched"]["heatcool"] == n - 2
    assert r["matched"]["fluxes"][0] == n - 4 and r["matched"]["fm"] == n - 4
    nr = CFG.nx[0]
    pre = r["inputs"].pre
    for q in far:
        j, i = divmod(q, nr)
        assert (pre[:, j, i] == 1.0).all()
        assert r["inputs"].xi[j, i] == 0.0            # the caller's (zero) array
        assert (r["fluxes"][0][:, 0, j, i] == 0.0).all()
    for q in near:
        j, i = divmod(q, nr)
        assert np.array_equal(pre[:, j, i], inp.pre[:, j, i])
        assert (r["fluxes"][0][:, 0, j, i] == 0.0).all()


def test_malformed_files_are_errors(pmx, tmp_path):
    """A record with fewer words than the format (trailing words are ignored
    by the reference's sscanf as well), a theta-flux file whose bin count
    differs from the r file's, and a fit record cut short are errors."""
    V = state()
    inp = setups.sirocco_inputs(CFG, V, fm_points=4)
    for name, mangle in (
            ("prefactors.dat", lambda t: t[: t.rfind(" ")] + "\n"),        # 10 words
            ("py_heatcool.dat", lambda t: t[: t.rfind(" ")] + "\n"),       # 17 words
            ("directional_flux_theta.dat", lambda t: t.replace("nangles 36", "nangles 35")),
            ("M_UV_data.dat", lambda t: t.rsplit(" ", 3)[0] + "\n")):      # short record
        d = tmp_path / name.replace(".", "_")
        d.mkdir()
        write_all(d, V, inp)
        f = d / name
        f.write_text(mangle(f.read_text()))
        with pytest.raises(pmx.PmxError):
            pmx.read_sirocco_files(CFG, d, V)


def test_negative_prefactors_are_counted(pmx, tmp_path):
    V = state()
    inp = setups.sirocco_inputs(CFG, V, heatcool=False)
    inp.pre[2, 3, 4] = -0.5
    inp.pre[5, 0, 0] = -1.0
    setups.write_sirocco_files(CFG, tmp_path, inp, V)
    r = pmx.read_sirocco_files(CFG, tmp_path, V)
    assert r["negative_prefactors"] == 2
    assert np.array_equal(r["inputs"].pre, inp.pre)


def test_readers_agree_with_the_reference_reader(pmx, tmp_path):
    """End to end against the compiled reference: the reference reads a
    set of SIROCCO files as a restart does, the oracle runs on this
    library's reading of the same files, and the results are identical.
    Some records are moved off their cell within the tolerances (all
    files) and some beyond them (prefactors.dat only: the reference leaves
    the cells other files do not name uninitialised, prefactors at 1)."""
    if not refrun.reference_binary("wind_cvidl_prod").exists():
        pytest.skip("reference build (make -C oracle/ref) not present")
    refrun.build(with_reference=False)
    cfg = dataclasses.replace(setups.cvidl_wind(32, 16, dfloor=1.e-15), first_dt=1.e-3,
                              shock_flattening=True, char_limiting=True, entropy_switch=True,
                              solver=setups.HLL, krad=999.0, alpharad=999.0)
    V = setups.wind_ic_mixed(cfg, rho0_cgs=1.e-11, t_range=(1.e2, 3.e7))
    inp = setups.sirocco_inputs(cfg, V, fm_points=12, fm_range=(1e-4, 1e-2))
    setups.write_flux_files(cfg, tmp_path)
    setups.write_sirocco_files(cfg, tmp_path, inp, V)
    for name, col, frac in (("prefactors.dat", 1, 5e-6), ("py_heatcool.dat", 2, 5e-6),
                            ("M_UV_data.dat", 2, 5e-7), ("directional_flux_theta.dat", 3, 5e-7)):
        _shift_records(tmp_path / name, frac, [7, 100], col)
    _shift_records(tmp_path / "prefactors.dat", 3e-5, [9, 200], 1)
    r = pmx.read_sirocco_files(cfg, tmp_path, V)
    n = cfg.nx[0] * cfg.nx[1]
    assert r["matched"]["prefactors"] == n - 2 and r["matched"]["fm"] == n
    assert r["matched"]["heatcool"] == n and r["matched"]["fluxes"] == [n, n, n]
    got = r["inputs"]
    for q in (9, 200):
        j, i = divmod(q, cfg.nx[0])
        assert (got.pre[:, j, i] == 1.0).all() and (inp.pre[:, j, i] != 1.0).all()
    ref = refrun.run_reference(cfg, V, 6, files_dir=tmp_path, t0=10.0)
    orc = refrun.run_wind_oracle(cfg, V, 6, t0=10.0, fluxes=r["fluxes"], sirocco=got)
    assert np.array_equal(orc["V"], ref["V"][:6])
    assert np.array_equal(orc["dt_seq"], ref["dt_seq"])
    # and the records beyond the tolerance matter
    orc2 = refrun.run_wind_oracle(cfg, V, 6, t0=10.0, fluxes=r["fluxes"],
                                  sirocco=dataclasses.replace(got, pre=inp.pre))
    assert not np.array_equal(orc2["V"], ref["V"][:6])
