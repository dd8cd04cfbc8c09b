"""GPU parity tests of the line-driven wind path (libplutomx pmx_wind_*).

Tolerances.  Everything in the wind step that uses only +, -, *, / and
sqrt is bit-identical to the reference; the only differences come from
the force multiplier's power M = k t^alpha of the line force (evaluated
from shared double-double logarithms, within about 0.6 ulp of the exact
power where the host libm's is within 0.5: the two differ in the last bit
for about 1.4% of arguments) and from the device math library's exp / log10
/ cosh in the cooling rates.  So:

* with k = 0 (no force multiplier), no cooling and no internal-boundary
  floors (whose pow(cs, 2) the device forms as cs*cs) the GPU step must be
  bit-identical to the oracle;
* otherwise every variable must satisfy |gpu - ref| <= 1e-12 |ref| +
  1e-12 max|ref| (relative error, with the variable's own scale as the
  floor for components that pass through zero, such as v_r in a wind that
  starts at rest), and the dt sequence must agree to 1e-12.
"""
import math

import numpy as np
import pytest

import golden_io
import setups
from oracle import refrun

pytestmark = pytest.mark.gpu

RTOL = 1e-12


@pytest.fixture(scope="module")
def pmx():
    import plutomx
    plutomx.load_library()
    return plutomx


def assert_close(a, b, what, rtol=RTOL):
    for nv in range(b.shape[0]):
        x, y = a[nv], b[nv]
        scale = np.max(np.abs(y))
        bad = np.abs(x - y) > rtol * np.abs(y) + rtol * scale
        assert not bad.any(), (what, nv, np.max(np.abs(x - y) / np.maximum(np.abs(y), 1e-300)),
                               int(bad.sum()))


def gpu_run(pmx, cfg, V0, nsteps, fluxes=None, grid=None, stepwise=False, sirocco=None,
            t0=0.0):
    with pmx.Wind(cfg, fluxes=fluxes, grid=grid, sirocco=sirocco) as w:
        w.set_prim(V0)
        if sirocco is not None:
            w.set_sirocco(sirocco)
        w.set_time(t0, cfg.first_dt, 0)
        dts = []
        for _ in range(nsteps):
            dts.append(w.get_time()["dt"])
            w.run(1)
        w.sync()
        return w.get_prim(), w.get_cons(), np.array(dts), w.get_time()


def fixture_inputs(g):
    grid = ({"x": g["x1"], "xl": g["x1l"], "xr": g["x1r"], "dx": g["dx1"]},
            {"x": g["x2"], "xl": g["x2l"], "xr": g["x2r"], "dx": g["dx2"]})
    return (g["flux_r"], g["flux_t"], g["flux_p"]), grid


@pytest.mark.parametrize("name", golden_io.names("wind"))
def test_wind_gpu_matches_reference_fixture(pmx, name):
    cfg, g = golden_io.load(name)
    fluxes, grid = fixture_inputs(g)
    sir, t0 = golden_io.sirocco(g)
    V, U, dts, tm = gpu_run(pmx, cfg, g["V0"], int(g["nsteps"]), fluxes, grid, sirocco=sir,
                            t0=t0)
    assert_close(V, g["V"], name + " V")
    assert_close(U, g["U"], name + " U")
    np.testing.assert_allclose(dts, g["dt_seq"], rtol=RTOL, atol=0)
    assert abs(tm["t"] - float(g["t"])) <= RTOL * float(g["t"])
    assert abs(tm["dt"] - float(g["dt"])) <= RTOL * float(g["dt"])


def test_wind_bitwise_without_force_multiplier(pmx):
    """k = 0 removes the only libm call from the hydro + line-force step:
    the GPU must then reproduce the oracle bit for bit (mixed state with
    floors, 64 x 32 grid, 10 steps)."""
    import dataclasses
    cfg = setups.WindConfig(nx=(64, 32), cooling=setups.COOL_NONE, first_dt=1e-3)
    cfg = dataclasses.replace(cfg, krad=0.0)
    V0 = setups.wind_ic_mixed(cfg, rho0_cgs=1e-11)
    o = refrun.run_wind_oracle(cfg, V0, 10)
    V, U, dts, tm = gpu_run(pmx, cfg, V0, 10)
    assert np.array_equal(V, o["V"]), np.abs(V - o["V"]).max()
    assert np.array_equal(U, o["U"]), np.abs(U - o["U"]).max()
    assert np.array_equal(dts, o["dt_seq"]) and tm["dt"] == o["dt"]


def test_reciprocal_division_is_bit_identical(pmx):
    """The line force divides by c, by the run's unit acceleration and by
    each cell's radial corner spacing through a precomputed reciprocal and
    two FMA corrections (where the run's fluxes, units and radii keep every
    operand in range).  On the device that evaluation must equal the IEEE
    quotient bit for bit: x of both signs spread over 2^-300..2^300, signed
    zeros, exact multiples of c and their neighbours, and values near the
    ends of the range."""
    rng = np.random.default_rng(7)
    n = 1 << 21
    m = rng.uniform(1.0, 2.0, n) * np.exp2(rng.integers(-300, 301, n).astype(float))
    x = np.where(rng.random(n) < 0.5, -m, m)
    x[:500] = 0.0
    x[500:1000] = -0.0
    cfg = setups.WindConfig(nx=(64, 32))
    ua = cfg.unit_velocity ** 2 / cfg.unit_length
    gr, _ = setups.wind_grid(cfg)
    xl = gr["x"] * cfg.unit_length
    dx0 = (xl[2:] - xl[:-2]) / 2.0              # corner spacing of each cell
    consts = [2.99792458e10, ua, 1.0 / 3.0, 3.0, 0.1] + list(rng.choice(dx0, 4))
    for c in consts:
        assert 2.0 ** -45 < c < 2.0 ** 45
        k = rng.integers(1, 1 << 40, 1 << 16).astype(float)
        mult = k * c
        e = rng.uniform(1.0, 2.0, 4096)
        edges = np.concatenate([e * 2.0 ** 850, e * 2.0 ** -850, c * e * 2.0 ** 850,
                                c * e * 2.0 ** -850])
        edges = np.concatenate([edges, -edges])
        xs = np.concatenate([x, mult, np.nextafter(mult, np.inf), np.nextafter(mult, -np.inf),
                             edges])
        assert pmx.check_div_rc(xs, c) == 0, c


def test_force_power_against_exact(pmx):
    """The line force's power pow(num / dvds, alpha) as the device evaluates
    it (log2 num and log2 dvds in double-double, shared between the calls,
    the quotient's rounding error folded in, one rounding) against the
    exact power of the rounded quotient (mpmath at 160 bits): within 0.6
    ulp (0.575 measured), within one ulp of the host libm's pow and equal
    to it for at least 97% of the arguments (98.6% measured),
    over the ranges of the wind (num = sigma_e rho v_th from 1e-17 to 0.1,
    dv/ds from 1e-12 to 1e4 s^-1, alpha 0.2 to 0.95)."""
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.prec = 160
    rng = np.random.default_rng(11)
    n = 4000
    num = 10.0 ** rng.uniform(-17, -1, n)
    dvds = 10.0 ** rng.uniform(-12, 4, n)
    for alpha in (0.6, 0.2 + 0.75 * rng.random()):
        out = pmx.check_force_power(num, dvds, 1.0, alpha)
        t = num / dvds                      # the reference's rounded quotient
        exact = np.array([float(mpmath.power(mpmath.mpf(x), mpmath.mpf(alpha))) for x in t])
        err = np.array([float((mpmath.mpf(o) - mpmath.power(mpmath.mpf(x), mpmath.mpf(alpha)))
                              / mpmath.mpf(float(np.spacing(e))))
                        for o, x, e in zip(out, t, exact)])
        assert np.abs(err).max() <= 0.6, (alpha, np.abs(err).max())
        # the host libm's pow (numpy's own power may take a SIMD
        # implementation that is not the libm's)
        lib = np.array([math.pow(x, alpha) for x in t])
        assert np.all(np.abs(out - lib) <= np.spacing(lib)), alpha
        assert np.mean(out == lib) >= 0.97, (alpha, np.mean(out == lib))
    # krad scales the rounded power (one more rounding, as the reference's
    # krad * pow)
    out = pmx.check_force_power(num, dvds, 0.2, 0.6)
    lib = 0.2 * np.array([math.pow(x, 0.6) for x in num / dvds])
    assert np.all(np.abs(out - lib) <= 2 * np.spacing(out))


def test_force_power_special_cases(pmx):
    """Zero, subnormal, overflowing and exact cases of the line force's
    power, and dv/ds <= 0 (the bin has M = 0)."""
    tiny = np.finfo(float).tiny
    num = np.array([0.0, 1e-310, 1e-300, 1.0, 1e-3, 2.0 ** -40, 1e-5, 1e-5, 1.0, np.nan, 1e-300])
    dvds = np.array([1.0, 1.0, 1e10, 1e-310, 1e300, 2.0 ** -20, 0.0, -3.0, 2.0 ** 10, 1.0,
                     1e-310])
    out = pmx.check_force_power(num, dvds, 1.0, 0.5)
    with np.errstate(over="ignore", divide="ignore", invalid="ignore"):
        t = num / dvds
    lib = np.array([math.pow(x, 0.5) if d > 0.0 else 0.0 for x, d in zip(t, dvds)])
    assert out[0] == 0.0                                   # num = 0
    assert t[1] < tiny and abs(out[1] - lib[1]) <= np.spacing(lib[1])   # subnormal num
    assert t[2] < tiny and abs(out[2] - lib[2]) <= np.spacing(lib[2])   # subnormal t
    assert out[3] == np.inf                                # t overflows to inf
    assert abs(out[4] - lib[4]) <= np.spacing(lib[4])      # t near the bottom of the normals
    assert out[5] == 2.0 ** -10                            # exact: (2^-20)^0.5
    assert out[6] == 0.0 and out[7] == 0.0                 # dv/ds <= 0
    assert out[8] == 2.0 ** -5                             # exact: (2^-10)^0.5
    assert np.isnan(out[9])
    assert abs(out[10] - lib[10]) <= np.spacing(lib[10])   # subnormal dv/ds, t = 1e10


def test_reciprocal_division_is_chosen_per_run(pmx):
    """A run whose fluxes hold a value outside 2^-600..2^600 (its line
    force then takes the division) matches the oracle bit for bit, like the
    runs that take the reciprocal."""
    import dataclasses
    cfg = dataclasses.replace(setups.WindConfig(nx=(32, 16), cooling=setups.COOL_NONE,
                                                first_dt=1e-3), krad=0.0)
    V0 = setups.wind_ic_mixed(cfg, rho0_cgs=1e-11)
    fr, ft, fp = (np.array(f, copy=True) for f in setups.wind_fluxes(cfg))
    fr[3, 0, 5, 7] = 1e-250
    o = refrun.run_wind_oracle(cfg, V0, 4, fluxes=(fr, ft, fp))
    V, U, dts, tm = gpu_run(pmx, cfg, V0, 4, fluxes=(fr, ft, fp))
    assert np.array_equal(V, o["V"]) and np.array_equal(U, o["U"])


@pytest.mark.parametrize("cooling", [setups.COOL_BLONDIN, setups.COOL_PLAIN])
def test_wind_c3_grid_matches_oracle(pmx, cooling):
    """The 512 x 256 wind grid of the benchmark workload, 4 steps."""
    cfg = setups.c3_wind(512, 256, cooling)
    V0 = setups.wind_ic(cfg)
    o = refrun.run_wind_oracle(cfg, V0, 4)
    V, U, dts, tm = gpu_run(pmx, cfg, V0, 4)
    assert_close(V, o["V"], "V")
    assert_close(U, o["U"], "U")
    np.testing.assert_allclose(dts, o["dt_seq"], rtol=RTOL, atol=0)


def test_wind_production_boundaries_match_oracle(pmx):
    """The production problem's internal-boundary floors and userdef sides
    on a 256 x 128 grid whose mixed state puts ~5% of the cells below the
    density floor and some below 500 K, 6 steps with plain cooling."""
    import dataclasses
    cfg = dataclasses.replace(setups.cvidl_wind(256, 128, dfloor=1.e-15), first_dt=1.e-3)
    V0 = setups.wind_ic_mixed(cfg, rho0_cgs=1.e-11, t_range=(1.e2, 3.e7))
    assert np.mean(V0[0] * cfg.unit_density < cfg.dfloor) > 0.02
    o = refrun.run_wind_oracle(cfg, V0, 6)
    assert refrun.wind_counters()["floor_cells"] > 0
    V, U, dts, tm = gpu_run(pmx, cfg, V0, 6)
    assert_close(V, o["V"], "V")
    assert_close(U, o["U"], "U")
    np.testing.assert_allclose(dts, o["dt_seq"], rtol=RTOL, atol=0)


def test_wind_shock_flattening_matches_oracle(pmx):
    """SHOCK_FLATTENING MULTID with the production boundaries on a 256 x 128
    grid whose mixed state has many shocked zones, 6 steps."""
    import dataclasses
    cfg = dataclasses.replace(setups.cvidl_wind(256, 128, dfloor=1.e-15), first_dt=1.e-3,
                              shock_flattening=True)
    V0 = setups.wind_ic_mixed(cfg, rho0_cgs=1.e-11, t_range=(1.e2, 3.e7))
    o = refrun.run_wind_oracle(cfg, V0, 6)
    assert refrun.wind_counters()["shock_zones"] > 1000
    V, U, dts, tm = gpu_run(pmx, cfg, V0, 6)
    assert_close(V, o["V"], "V")
    assert_close(U, o["U"], "U")
    np.testing.assert_allclose(dts, o["dt_seq"], rtol=RTOL, atol=0)


def test_wind_char_limiting_matches_oracle(pmx):
    """The production problem's reconstruction (CHAR_LIMITING YES with
    SHOCK_FLATTENING MULTID) and boundaries on a 256 x 128 mixed state,
    6 steps."""
    import dataclasses
    cfg = dataclasses.replace(setups.cvidl_wind(256, 128, dfloor=1.e-15), first_dt=1.e-3,
                              shock_flattening=True, char_limiting=True)
    V0 = setups.wind_ic_mixed(cfg, rho0_cgs=1.e-11, t_range=(1.e2, 3.e7))
    o = refrun.run_wind_oracle(cfg, V0, 6)
    V, U, dts, tm = gpu_run(pmx, cfg, V0, 6)
    assert_close(V, o["V"], "V")
    assert_close(U, o["U"], "U")
    np.testing.assert_allclose(dts, o["dt_seq"], rtol=RTOL, atol=0)


def test_wind_production_reconstruction_matches_oracle(pmx):
    """The production problem's full reconstruction (CHAR_LIMITING YES,
    SHOCK_FLATTENING MULTID, ENTROPY_SWITCH ALWAYS) and boundaries on a
    256 x 128 mixed state, 6 steps."""
    import dataclasses
    cfg = dataclasses.replace(setups.cvidl_wind(256, 128, dfloor=1.e-15), first_dt=1.e-3,
                              shock_flattening=True, char_limiting=True, entropy_switch=True)
    V0 = setups.wind_ic_mixed(cfg, rho0_cgs=1.e-11, t_range=(1.e2, 3.e7))
    o = refrun.run_wind_oracle(cfg, V0, 6)
    V, U, dts, tm = gpu_run(pmx, cfg, V0, 6)
    assert_close(V, o["V"], "V")
    assert_close(U, o["U"], "U")
    np.testing.assert_allclose(dts, o["dt_seq"], rtol=RTOL, atol=0)


def test_wind_production_setup_with_hll_matches_oracle(pmx):
    """The production problem's set-up with its HLL solver, 128 x 64,
    6 steps."""
    import dataclasses
    cfg = dataclasses.replace(setups.cvidl_wind(128, 64, dfloor=1.e-15), first_dt=1.e-3,
                              shock_flattening=True, char_limiting=True, entropy_switch=True,
                              solver=setups.HLL)
    V0 = setups.wind_ic_mixed(cfg, rho0_cgs=1.e-11, t_range=(1.e2, 3.e7))
    o = refrun.run_wind_oracle(cfg, V0, 6)
    V, U, dts, tm = gpu_run(pmx, cfg, V0, 6)
    assert_close(V, o["V"], "V")
    assert_close(U, o["U"], "U")


def test_wind_entropy_switch_without_char_limiting_matches_oracle(pmx):
    """The entropy switch with van Leer / minmod scalars (no characteristic
    limiting) and shock flattening, 128 x 64, 6 steps."""
    import dataclasses
    cfg = dataclasses.replace(setups.cvidl_wind(128, 64, dfloor=1.e-15), first_dt=1.e-3,
                              shock_flattening=True, entropy_switch=True)
    V0 = setups.wind_ic_mixed(cfg, rho0_cgs=1.e-11, t_range=(1.e2, 3.e7))
    o = refrun.run_wind_oracle(cfg, V0, 6)
    V, U, dts, tm = gpu_run(pmx, cfg, V0, 6)
    assert_close(V, o["V"], "V")
    assert_close(U, o["U"], "U")


def test_wind_dt_too_small_stop_is_reported(pmx):
    """NextTimeStep's stop (dt < 1e-9 first_dt) sets PMX_FLAG_DT_TOO_SMALL
    for the wind too; a normal run leaves it clear."""
    import dataclasses
    base = setups.WindConfig(nx=(32, 16), cooling=setups.COOL_PLAIN, first_dt=1.e-3)
    V0 = setups.wind_ic(base)
    for cfg, expect in ((base, False), (dataclasses.replace(base, first_dt=1.e9), True)):
        with pmx.Wind(cfg) as w:
            w.set_prim(V0)
            w.set_time(0.0, base.first_dt, 0)
            w.run(4)
            w.sync()
            tm = w.get_time()
            assert bool(tm["flags"] & pmx.PMX_FLAG_DT_TOO_SMALL) == expect
            # the reference quits at the first NextTimeStep (after step 1,
            # with cooling): the loop halts there
            assert tm["step"] == (2 if expect else 4)


@pytest.mark.parametrize("name", [n for n in golden_io.names("wind") if "_tstop_" in n])
def test_wind_tstop_cut_and_halt(pmx, name):
    """pluto.ini tstop for the wind's main loop (Src/main.c:226-229): the
    reference fixtures end at an odd and an even step (the Strang order and
    the every-second-step NextTimeStep depend on the parity); asked for more
    steps, the eager loop and the graph replay both stop there with
    PMX_FLAG_TSTOP and report the steps actually taken."""
    cfg, g = golden_io.load(name)
    fluxes, grid = fixture_inputs(g)
    n = int(g["nsteps"])
    for graph in (False, True):
        with pmx.Wind(cfg, fluxes=fluxes, grid=grid) as w:
            w.set_prim(g["V0"])
            w.set_time(0.0, cfg.first_dt, 0)
            w.run(n + 5, graph=graph)
            w.sync()
            tm = w.get_time()
            assert tm["flags"] & pmx.PMX_FLAG_TSTOP and tm["step"] == n, (graph, tm)
            assert_close(w.get_prim(), g["V"], f"{name} V graph={graph}")
            assert_close(w.get_cons(), g["U"], f"{name} U graph={graph}")
            assert tm["t"] == float(g["tstop"]) or abs(tm["t"] - float(g["t"])) <= RTOL * tm["t"]
            assert abs(tm["dt"] - float(g["dt"])) <= RTOL * float(g["dt"])


@pytest.mark.parametrize("cooling", [setups.COOL_BLONDIN, setups.COOL_PLAIN])
@pytest.mark.parametrize("t0", [0.0, 5.0])
def test_wind_single_cooling_call(pmx, cooling, t0):
    """SplitSource alone on a mixed state (stiff cells, sub-1e4 K cells,
    both tracer branches), before and after the g_time = 3 switch to the
    SIROCCO xi / T_r arrays; dt_cool of the plain integrator too."""
    cfg = setups.WindConfig(nx=(64, 32), cooling=cooling)
    V0 = setups.wind_ic_mixed(cfg, rho0_cgs=1e-10)
    o = refrun.run_wind_cool_oracle(cfg, V0, 1e-2, t0)
    with pmx.Wind(cfg) as w:
        w.set_prim(V0)
        w.set_time(t0, 1e-2, 0)
        w.split_source()
        w.sync()
        V, tm = w.get_prim(), w.get_time()
    assert np.array_equal(V[[0, 1, 2, 3, 5]], o["V"][[0, 1, 2, 3, 5]])
    np.testing.assert_allclose(V[4], o["V"][4], rtol=RTOL, atol=0)
    if cooling == setups.COOL_PLAIN:
        assert abs(tm["dt_cool"] - o["dt_cool"]) <= RTOL * o["dt_cool"]
    assert tm["flags"] & 256 == 0


def test_wind_granular_calls_match_run(pmx):
    """advance_step / split_source in the reference's Strang order give the
    same state as run() (which also keeps dt on the device)."""
    cfg = setups.WindConfig(nx=(48, 24), cooling=setups.COOL_PLAIN)
    V0 = setups.wind_ic_mixed(cfg, rho0_cgs=1e-11)
    with pmx.Wind(cfg) as w:
        w.set_prim(V0)
        w.set_time(0.0, 1e-3, 0)
        w.run(2)
        w.sync()
        Va = w.get_prim()
    with pmx.Wind(cfg) as w:
        w.set_prim(V0)
        w.set_time(0.0, 1e-3, 0)
        w.advance_step()
        w.split_source()
        w.sync()
        w.set_time(1e-3, 1e-3, 1)
        w.split_source()
        w.advance_step()
        w.sync()
        Vb = w.get_prim()
    assert np.array_equal(Va, Vb)


def test_wind_run_is_deterministic(pmx):
    cfg = setups.c3_wind(512, 256, setups.COOL_PLAIN)
    V0 = setups.wind_ic_mixed(cfg, rho0_cgs=1e-11)
    out = []
    for _ in range(2):
        V, U, dts, tm = gpu_run(pmx, cfg, V0, 6)
        out.append((V, U, dts))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])


@pytest.mark.parametrize("dt", [1e-2, 1.0, 30.0])
def test_brem_only_cooling_closed_forms_on_device(pmx, dt):
    """Bremsstrahlung only (prefactors 0 except brem_c): the device's
    sub-cycled integrator follows sqrt(E) = sqrt(E0) - B t / 2 to 1e-4, and
    its BLONDIN step lands within 1 K of the trapezoid root
    (closed forms in tests/test_cooling_closed_form.py)."""
    import test_cooling_closed_form as cf
    for cooling in (setups.COOL_PLAIN, setups.COOL_BLONDIN):
        cfg, V, pre, kelvin = cf.brem_case(cooling)
        with pmx.Wind(cfg) as w:
            w.set_prim(V)
            w.set_prefactors(pre)
            w.set_time(10.0, dt, 0)
            w.split_source()
            w.sync()
            p = w.get_prim()[setups.PRS, 0]
        if cooling == setups.COOL_PLAIN:
            np.testing.assert_allclose(p, cf.analytic_plain(cfg, V, dt, kelvin), rtol=1e-4)
        else:
            p_an, Tf = cf.analytic_blondin(cfg, V, dt)
            assert (np.abs(p / p_an - 1.0) * Tf).max() <= 1.0


@pytest.mark.parametrize("kind", ["plain", "production"])
def test_wind_graph_replay_matches_eager(pmx, kind):
    """pmx_wind_run_graph replays 8 captured steps from an even step: every
    replay, with host reads and synchronisations between them, matches the
    eager run bit for bit, from an odd start step (one eager step first)
    and with a remainder that runs eagerly."""
    import dataclasses
    if kind == "plain":
        cfg = setups.WindConfig(nx=(64, 32), first_dt=1e-3)
    else:
        cfg = dataclasses.replace(setups.cvidl_wind(64, 32, dfloor=1.e-15), first_dt=1.e-3,
                                  shock_flattening=True, char_limiting=True,
                                  entropy_switch=True, solver=setups.HLL)
    V0 = setups.wind_ic_mixed(cfg, rho0_cgs=1.e-11, t_range=(1.e2, 3.e7))
    eager = []
    with pmx.Wind(cfg) as w:
        w.set_prim(V0)
        w.set_time(0.0, cfg.first_dt, 0)
        w.run(1)
        for _ in range(3):
            w.run(9)
            w.sync()
            eager.append((w.get_prim(), w.get_time()))
    with pmx.Wind(cfg) as w:
        w.set_prim(V0)
        w.set_time(0.0, cfg.first_dt, 0)
        w.run(1, graph=True)
        for q in range(3):
            w.run(9, graph=True)     # steps 1..9: 1 eager + 8; 10..18: 8 + 1; 19..27: 1 + 8
            w.sync()
            V, tm = w.get_prim(), w.get_time()
            assert np.array_equal(V, eager[q][0]), q
            assert (tm["t"], tm["dt"], tm["step"]) == (eager[q][1]["t"], eager[q][1]["dt"],
                                                       eager[q][1]["step"])
