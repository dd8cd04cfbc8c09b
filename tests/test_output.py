"""CPU tests of the output path: the production problem's user variables
(ComputeUserVar, Test_Problems/LineDrivenWind/cv_idl/userdef_output.c)
and the .dbl / dbl.out / restart.out writers (Src/write_data.c,
Src/restart.c).

The fixtures wind_output_*_32x16 hold what the compiled reference wrote
with the production uservar list and a .dbl file every step: the last
data.NNNN.dbl (V and the 23 user variables after 6 steps), dbl.out and the
fields of every restart.out record.  The oracle must reproduce the user
variables bit for bit; the writers must reproduce the files byte for byte
(restart.out up to the 40 fill bytes, which the reference leaves unset).
"""
import numpy as np
import pytest

import golden_io
import setups
from oracle import refrun

NAMES = ["wind_output_plain_32x16", "wind_output_prod_32x16"]


@pytest.fixture(scope="module", autouse=True)
def _build():
    refrun.build(with_reference=False)


def oracle_run(name):
    cfg, g = golden_io.load(name)
    sir, t0 = golden_io.sirocco(g)
    grid = ({"x": g["x1"], "xl": g["x1l"], "xr": g["x1r"], "dx": g["dx1"]},
            {"x": g["x2"], "xl": g["x2l"], "xr": g["x2r"], "dx": g["dx2"]})
    r = refrun.run_wind_oracle(cfg, g["V0"], int(g["nsteps"]), t0=t0, sirocco=sir,
                               fluxes=(g["flux_r"], g["flux_t"], g["flux_p"]), grid=grid,
                               user_vars=True)
    return cfg, g, sir, r


@pytest.mark.parametrize("name", NAMES)
def test_user_vars_match_reference_dbl(name):
    cfg, g, _, r = oracle_run(name)
    dbl = g["dbl_final"]
    assert np.array_equal(r["V"][:, 0], dbl[:6])
    for q, nm in enumerate(setups.USER_VARS):
        assert np.array_equal(r["uv"][q], dbl[6 + q]), nm


def test_user_var_layout_against_inputs():
    """The fixtures' user variables sit where USER_VARS says: before t = 1
    XI is L_x / (n_H r^2) with n_H the nh column and T_r is T_x; after a
    restart XI / T_r are the SIROCCO arrays after the reader's floors and
    the *_pre columns are the prefactors in prefactors.dat order
    (comp_h, comp_c, xray_h, brem_c, line_c, xi_ion); gp is 0."""
    uv = {n: q for q, n in enumerate(setups.USER_VARS)}
    cfg, g = golden_io.load("wind_output_plain_32x16")
    d = g["dbl_final"][6:]
    gr, _ = setups.wind_grid(cfg)
    r = gr["x"][2:-2][None, :] * cfg.unit_length
    np.testing.assert_allclose(d[uv["XI"]], cfg.l_star * cfg.f_x / (d[uv["nh"]] * r * r),
                               rtol=1e-15)
    assert (d[uv["T_r"]] == cfg.t_x).all() and (d[uv["gp"]] == 0.0).all()
    assert (d[uv["ch_pre"]] == 1.0).all()
    cfg, g = golden_io.load("wind_output_prod_32x16")
    sir, _ = golden_io.sirocco(g)
    d = g["dbl_final"][6:]
    xi, t_r = sir.heatcool_floored()
    assert np.array_equal(d[uv["XI"]], xi) and np.array_equal(d[uv["T_r"]], t_r)
    for nm, plane in (("ch_pre", 0), ("cc_pre", 1), ("xh_pre", 2), ("bc_pre", 3),
                      ("lc_pre", 4), ("XI_pre", 5)):
        assert np.array_equal(d[uv[nm]], sir.pre[plane]), nm
    # after t = 3 the rates carry the prefactors: ch / ch_pre is the bare rate
    bare = 8.9e-36 * d[uv["XI"]] * cfg.t_x
    np.testing.assert_allclose(d[uv["ch"]], sir.pre[0] * bare, rtol=1e-15)
    # the force multipliers are capped at 4400 and M_max1 is the maximum
    assert (d[uv["M_max1"]] <= 4400.0).all() and (d[uv["M_max2"]] <= d[uv["M_max1"]]).all()


@pytest.fixture(scope="module")
def pmx():
    import plutomx
    plutomx.load_library()
    return plutomx


@pytest.mark.parametrize("name", NAMES)
def test_dbl_writer_matches_reference(pmx, tmp_path, name):
    """Writing the reference's outputs again (the final state and the
    times of every output, taken from its restart records) gives its
    data.0005.dbl and dbl.out byte for byte."""
    cfg, g = golden_io.load(name)
    names = list(pmx.PRIM_NAMES) + list(setups.USER_VARS)
    dbl = g["dbl_final"][:, None]          # [nvar][1][nt][nr]
    for nf, (st, t, dt) in enumerate(zip(g["rs_nstep"], g["rs_t"], g["rs_dt"])):
        pmx.write_dbl(tmp_path, nf, dbl, names, t, dt, int(st))
    last = len(g["rs_nstep"]) - 1
    assert (tmp_path / f"data.{last:04d}.dbl").read_bytes() == g["dbl_final"].tobytes()
    assert (tmp_path / "dbl.out").read_text() == str(g["dbl_out"])


@pytest.mark.parametrize("name", NAMES)
def test_restart_writer_matches_reference(pmx, tmp_path, name):
    cfg, g = golden_io.load(name)
    for c, (st, nf, t, dt) in enumerate(zip(g["rs_nstep"], g["rs_nfile"], g["rs_t"],
                                            g["rs_dt"])):
        pmx.write_restart(tmp_path, c, int(st), nf, t, dt)
    b = (tmp_path / "restart.out").read_bytes()
    assert len(b) == 128 * len(g["rs_nstep"])
    rec = np.dtype([("nstep", "<i4"), ("nfile", "<i4", 16), ("pad", "<i4"), ("t", "<f8"),
                    ("dt", "<f8"), ("fill", "V40")])
    r = np.frombuffer(b, dtype=rec)
    assert np.array_equal(r["nstep"], g["rs_nstep"]) and np.array_equal(r["nfile"], g["rs_nfile"])
    assert np.array_equal(r["t"], g["rs_t"]) and np.array_equal(r["dt"], g["rs_dt"])
    assert (r["pad"] == 0).all()


def test_dbl_out_rewrites_from_nfile(pmx, tmp_path):
    """dbl.out keeps its first nfile lines and is rewritten from there (an
    output number written again replaces its line)."""
    data = np.zeros((1, 1, 2, 3))
    for nf in range(3):
        pmx.write_dbl(tmp_path, nf, data, ["rho"], 0.5 * nf, 0.25, nf)
    pmx.write_dbl(tmp_path, 1, data, ["rho"], 9.0, 0.25, 7)
    lines = (tmp_path / "dbl.out").read_text().splitlines()
    assert lines[0].startswith("0 0.000000e+00 2.500000e-01 0 single_file little rho")
    assert lines[1].startswith("1 9.000000e+00 2.500000e-01 7 single_file little rho")
    assert lines[2].startswith("2 1.000000e+00")


@pytest.mark.parametrize("name", NAMES)
def test_restart_readers_round_trip(pmx, tmp_path, name):
    """Files written by the writers read back through the restart readers:
    output 3's record (step, t, dt) and data."""
    cfg, g = golden_io.load(name)
    names = list(pmx.PRIM_NAMES) + list(setups.USER_VARS)
    for nf, (st, t, dt) in enumerate(zip(g["rs_nstep"], g["rs_t"], g["rs_dt"])):
        data = g["dbl_final"] if nf == 5 else np.concatenate([g["dbl_3"], g["dbl_final"][6:]])
        pmx.write_dbl(tmp_path, nf, data[:, None], names, t, dt, int(st))
        pmx.write_restart(tmp_path, nf, int(st), g["rs_nfile"][nf], t, dt)
    r = pmx.read_restart(tmp_path, 3, 6, (1, cfg.nx[1], cfg.nx[0]))
    assert r["nstep"] == g["rs_nstep"][3] and r["t"] == g["rs_t"][3] and r["dt"] == g["rs_dt"][3]
    assert r["swap"] == 0 and r["nfile"][0] == 3
    assert np.array_equal(r["data"][:, 0], g["dbl_3"])
    with pytest.raises(pmx.PmxError, match="does not exist"):
        pmx.read_restart(tmp_path, 9, 6, (1, cfg.nx[1], cfg.nx[0]))


def test_restart_readers_read_the_reference_files(pmx, tmp_path):
    """The readers on the files the compiled reference wrote itself (skipped
    where oracle/_ref is not built)."""
    name = "wind_output_plain_32x16"
    cfg, g = golden_io.load(name)
    if not refrun.reference_binary(cfg.reference_problem()).exists():
        pytest.skip("reference build (make -C oracle/ref) not present")
    r = refrun.run_reference(cfg, g["V0"], int(g["nsteps"]), output=True)
    d = tmp_path
    for q, b in enumerate(r["dbl"]):
        (d / f"data.{q:04d}.dbl").write_bytes(b)
    (d / "dbl.out").write_text(r["dbl_out"])
    (d / "restart.out").write_bytes(r["restart_out"])
    rr = pmx.read_restart(d, 3, 6, (1, cfg.nx[1], cfg.nx[0]))
    assert rr["nstep"] == g["rs_nstep"][3] and rr["t"] == g["rs_t"][3]
    assert np.array_equal(rr["data"][:, 0], g["dbl_3"])
