"""GPU tests of the output snapshot (pmx_wind_snapshot / _wait): the state
and the production problem's user variables computed on the device match
what the compiled reference wrote to its last .dbl file (fixtures
wind_output_*_32x16), and the snapshot is the state at the point it was
queued even when later steps run before it is collected.

Tolerance as for the wind step: the user variables pass through pow / exp /
log10 (force multiplier, line-cooling rate, n_e/n_H fit), which the device
library may round differently in the last bit: 1e-12 relative, with the
variable's scale as the floor.  gp is exactly 0."""
import numpy as np
import pytest

import golden_io
import setups

pytestmark = pytest.mark.gpu

RTOL = 1e-12


@pytest.fixture(scope="module")
def pmx():
    import plutomx
    plutomx.load_library()
    return plutomx


def assert_close(a, b, what):
    scale = np.max(np.abs(b))
    bad = np.abs(a - b) > RTOL * np.abs(b) + RTOL * scale
    assert not bad.any(), (what, int(bad.sum()),
                           np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def fixture_wind(pmx, name):
    cfg, g = golden_io.load(name)
    sir, t0 = golden_io.sirocco(g)
    grid = ({"x": g["x1"], "xl": g["x1l"], "xr": g["x1r"], "dx": g["dx1"]},
            {"x": g["x2"], "xl": g["x2l"], "xr": g["x2r"], "dx": g["dx2"]})
    w = pmx.Wind(cfg, fluxes=(g["flux_r"], g["flux_t"], g["flux_p"]), grid=grid, sirocco=sir)
    w.set_prim(g["V0"])
    if sir is not None:
        w.set_sirocco(sir)
    w.set_time(t0, cfg.first_dt, 0)
    return cfg, g, w


@pytest.mark.parametrize("name", ["wind_output_plain_32x16", "wind_output_prod_32x16"])
def test_snapshot_matches_reference_dbl(pmx, name):
    cfg, g, w = fixture_wind(pmx, name)
    with w:
        w.run(int(g["nsteps"]))
        w.snapshot(user_vars=True)
        s = w.snapshot_wait()
    dbl = g["dbl_final"]
    for nv in range(6):
        assert_close(s["V"][nv, 0], dbl[nv], f"V{nv}")
    for q, nm in enumerate(setups.USER_VARS):
        assert_close(s["uv"][q, 0], dbl[6 + q], nm)
    assert (s["uv"][setups.USER_VARS.index("gp")] == 0.0).all()
    assert abs(s["t"] - g["rs_t"][-1]) <= RTOL * abs(g["rs_t"][-1])
    assert abs(s["dt"] - g["rs_dt"][-1]) <= RTOL * g["rs_dt"][-1]


def test_snapshot_is_taken_when_queued(pmx):
    """Steps queued after the snapshot do not leak into it: the snapshot
    equals the state read synchronously at that point, while the context
    has moved on; a second snapshot reuses the staging buffer safely."""
    cfg, g, w = fixture_wind(pmx, "wind_output_plain_32x16")
    with w:
        w.run(3)
        V3 = w.get_prim()
        w.snapshot(user_vars=True)
        w.run(4)
        s = w.snapshot_wait()
        assert np.array_equal(s["V"], V3)
        V7 = w.get_prim()
        assert not np.array_equal(V7, V3)
        w.snapshot(user_vars=False)
        w.run(1)
        s2 = w.snapshot_wait(copy=False)
        assert np.array_equal(s2["V"], V7) and s2["uv"] is None
        assert not s2["V"].flags.writeable
        # the zero-copy view holds the buffer: no new snapshot over it ...
        with pytest.raises(pmx.PmxError, match="still referenced"):
            w.snapshot(user_vars=False)
        V = s2["V"]
        del s2
        with pytest.raises(pmx.PmxError, match="still referenced"):
            w.snapshot(user_vars=False)
        del V
        V8 = w.get_prim()
        w.snapshot(user_vars=False)
        w.run(1)
        s3 = w.snapshot_wait(copy=False)
    # ... and closing the context leaves the buffer to the view
    assert np.array_equal(s3["V"], V8)
    del s3


def test_user_vars_need_a_step(pmx):
    cfg, g, w = fixture_wind(pmx, "wind_output_plain_32x16")
    with w:
        with pytest.raises(pmx.PmxError, match="need a step"):
            w.snapshot(user_vars=True)
        w.snapshot(user_vars=False)
        assert np.array_equal(w.snapshot_wait()["V"], g["V0"])


@pytest.mark.parametrize("name", ["wind_output_plain_32x16", "wind_output_prod_32x16"])
def test_restart_from_output_continues_like_the_reference(pmx, tmp_path, name):
    """A run restarted from output 3 (read back with the restart readers
    from files the writers produced) reaches the reference's final output:
    the reference restarted from the same output reaches the state of its
    uninterrupted run, which the fixture holds."""
    cfg, g = golden_io.load(name)
    sir, _ = golden_io.sirocco(g)
    names = list(pmx.PRIM_NAMES)
    for nf, (st, t, dt) in enumerate(zip(g["rs_nstep"][:4], g["rs_t"][:4], g["rs_dt"][:4])):
        pmx.write_dbl(tmp_path, nf, g["dbl_3"][:, None], names, t, dt, int(st))
        pmx.write_restart(tmp_path, nf, int(st), g["rs_nfile"][nf], t, dt)
    r = pmx.read_restart(tmp_path, 3, 6, (1, cfg.nx[1], cfg.nx[0]))
    grid = ({"x": g["x1"], "xl": g["x1l"], "xr": g["x1r"], "dx": g["dx1"]},
            {"x": g["x2"], "xl": g["x2l"], "xr": g["x2r"], "dx": g["dx2"]})
    with pmx.Wind(cfg, fluxes=(g["flux_r"], g["flux_t"], g["flux_p"]), grid=grid,
                  sirocco=sir) as w:
        w.set_prim(r["data"])
        if sir is not None:
            w.set_sirocco(sir)
        w.set_time(r["t"], r["dt"], r["nstep"])
        w.run(int(g["rs_nstep"][-1]) - r["nstep"])
        w.snapshot(user_vars=True)
        s = w.snapshot_wait()
    dbl = g["dbl_final"]
    for nv in range(6):
        assert_close(s["V"][nv, 0], dbl[nv], f"V{nv}")
    for q, nm in enumerate(setups.USER_VARS):
        assert_close(s["uv"][q, 0], dbl[6 + q], nm)
