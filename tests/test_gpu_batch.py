"""GPU tests of the batched sweep (pmx_wind_batch_*): independent wind
grids of uneven sizes and parameters advanced together must give, grid by
grid, exactly the state each grid reaches when it is run on its own (the
arithmetic per cell is the same; only the scheduling differs, and the step
reductions are max / min, which do not depend on order)."""
import dataclasses

import numpy as np
import pytest

import setups

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pmx():
    import plutomx
    plutomx.load_library()
    return plutomx


def small_batch():
    base = [((48, 24), 1e38, 0.2), ((40, 32), 3e37, 0.35), ((64, 16), 1e37, 0.1)]
    cfgs = []
    for nx, ls, kr in base:
        cfg = setups.WindConfig(nx=nx, cooling=setups.COOL_PLAIN, first_dt=1e-3)
        cfgs.append(dataclasses.replace(cfg, l_star=ls, krad=kr))
    return cfgs


def states(cfgs, rho0=1e-11):
    return [setups.wind_ic_mixed(c, seed=30 + g, rho0_cgs=rho0) for g, c in enumerate(cfgs)]


def run_single(pmx, cfg, V0, nsteps, mlp=None, sirocco=None, t0=0.0):
    with pmx.Wind(cfg, sirocco=sirocco) as w:
        w.set_prim(V0)
        if sirocco is not None:
            w.set_sirocco(sirocco)
        w.set_time(t0, cfg.first_dt, 0)
        if mlp is not None:
            w.set_surrogate(mlp)
        w.run(nsteps)
        w.sync()
        return w.get_prim(), w.get_cons(), w.get_time()


def run_batch(pmx, cfgs, V0s, nsteps, graph=False, tune=False, mlp=None, sirocco=None,
              t0=0.0):
    sirocco = sirocco or [None] * len(cfgs)
    with pmx.WindBatch(cfgs, sirocco=sirocco) as b:
        for w, cfg, V0, sir in zip(b.grids, cfgs, V0s, sirocco):
            w.set_prim(V0)
            if sir is not None:
                w.set_sirocco(sir)
            w.set_time(t0, cfg.first_dt, 0)
        b.set_step(0)
        if mlp is not None:
            b.set_surrogate(mlp)
        if tune:
            b.autotune()
        b.run(nsteps, graph=graph)
        b.sync()
        return [(w.get_prim(), w.get_cons(), w.get_time()) for w in b.grids]


def test_batch_matches_single_grids(pmx):
    cfgs = small_batch()
    V0s = states(cfgs)
    batch = run_batch(pmx, cfgs, V0s, 6)
    for cfg, V0, (Vb, Ub, tb) in zip(cfgs, V0s, batch):
        Vs, Us, ts = run_single(pmx, cfg, V0, 6)
        assert np.array_equal(Vb, Vs) and np.array_equal(Ub, Us)
        assert tb["t"] == ts["t"] and tb["dt"] == ts["dt"]


def test_batch_autotuned_graph_replay_matches_eager(pmx):
    cfgs = small_batch()
    V0s = states(cfgs)
    eager = run_batch(pmx, cfgs, V0s, 7)
    tuned = run_batch(pmx, cfgs, V0s, 7, graph=True, tune=True)
    for (Ve, Ue, te), (Vt, Ut, tt) in zip(eager, tuned):
        assert np.array_equal(Ve, Vt) and np.array_equal(Ue, Ut)
        assert te["t"] == tt["t"] and te["dt"] == tt["dt"]


@pytest.mark.parametrize("shape", range(6))
def test_every_tile_shape(pmx, shape):
    cfgs = small_batch()[:2]
    V0s = states(cfgs)
    ref = run_batch(pmx, cfgs, V0s, 2)
    with pmx.WindBatch(cfgs) as b:
        for w, cfg, V0 in zip(b.grids, cfgs, V0s):
            w.set_prim(V0)
            w.set_time(0.0, cfg.first_dt, 0)
        b.set_step(0)
        b.set_shape(shape, 7)     # 7 persistent workgroups: every one loops
        b.run(2)
        b.sync()
        got = [w.get_prim() for w in b.grids]
    for (Vr, _, _), Vg in zip(ref, got):
        assert np.array_equal(Vr, Vg)


def test_batch_with_production_boundaries_matches_single_grids(pmx):
    """Grids with the production problem's floors and userdef sides (one
    also with shock flattening, one with flattening and characteristic
    limiting, one with its full reconstruction incl. the entropy switch) in
    a batch with two plain grids: every grid as on its own, bit for bit."""
    cv = dataclasses.replace(setups.cvidl_wind(40, 24, dfloor=1.e-15), first_dt=1.e-3)
    cvf = dataclasses.replace(setups.cvidl_wind(32, 16, dfloor=1.e-15), first_dt=1.e-3,
                              shock_flattening=True)
    cvc = dataclasses.replace(cvf, nx=(24, 24), char_limiting=True)
    cvp = dataclasses.replace(cvc, nx=(32, 24), entropy_switch=True)
    cfgs = small_batch()[:2] + [cv, cvf, cvc, cvp]
    V0s = [setups.wind_ic_mixed(c, seed=50 + g, rho0_cgs=1.e-11, t_range=(1.e2, 3.e7))
           for g, c in enumerate(cfgs)]
    batch = run_batch(pmx, cfgs, V0s, 6, graph=True)
    for cfg, V0, (Vb, Ub, tb) in zip(cfgs, V0s, batch):
        Vs, Us, ts = run_single(pmx, cfg, V0, 6)
        assert np.array_equal(Vb, Vs) and np.array_equal(Ub, Us)
        assert tb["dt"] == ts["dt"]


def test_batch_with_sirocco_restart_inputs_matches_single_grids(pmx):
    """A restarted cycle (t0 = 10: the cooling reads each grid's xi / T_r,
    with per-cell prefactors) over plain grids, a grid with the tabulated
    force multiplier and a production grid with it: every grid as on its
    own, bit for bit, with graph replay."""
    tab = setups.WindConfig(nx=(40, 24), krad=999.0, alpharad=999.0, first_dt=1e-3)
    prod = dataclasses.replace(setups.cvidl_wind(32, 24, dfloor=1.e-15), first_dt=1.e-3,
                               shock_flattening=True, char_limiting=True, entropy_switch=True,
                               solver=setups.HLL, krad=999.0, alpharad=999.0)
    cfgs = small_batch()[:2] + [tab, prod]
    V0s = [setups.wind_ic_mixed(c, seed=60 + g, rho0_cgs=1.e-11, t_range=(1.e2, 3.e7))
           for g, c in enumerate(cfgs)]
    sirs = [setups.sirocco_inputs(c, V, seed=70 + g, fm_points=10 if c.krad == 999.0 else 0,
                                  fm_range=(1e-4, 1e-2)) for g, (c, V) in enumerate(zip(cfgs, V0s))]
    batch = run_batch(pmx, cfgs, V0s, 6, graph=True, sirocco=sirs, t0=10.0)
    for cfg, V0, sir, (Vb, Ub, tb) in zip(cfgs, V0s, sirs, batch):
        Vs, Us, ts = run_single(pmx, cfg, V0, 6, sirocco=sir, t0=10.0)
        assert np.array_equal(Vb, Vs) and np.array_equal(Ub, Us)
        assert tb["dt"] == ts["dt"]


def test_batch_surrogate_matches_single_grids(pmx):
    """Stiff cells (dense gas, long dt) take the surrogate path; an
    untrained network still yields certified (or fallback) results, and the
    batch must agree with each grid's own surrogate run bit for bit."""
    cfgs = [dataclasses.replace(c, first_dt=3e-2) for c in small_batch()]
    V0s = states(cfgs, rho0=1e-10)
    with pmx.Mlp(seed=3) as m:
        batch = run_batch(pmx, cfgs, V0s, 4, graph=True, mlp=m)
        for cfg, V0, (Vb, Ub, tb) in zip(cfgs, V0s, batch):
            Vs, Us, ts = run_single(pmx, cfg, V0, 4, mlp=m)
            assert np.array_equal(Vb, Vs) and np.array_equal(Ub, Us)
            assert tb["dt"] == ts["dt"]


def test_batch_with_roe_grids_matches_single_grids(pmx):
    """Roe grids (one plain, one with the production boundaries and shock
    flattening, whose flagged zones take Roe's HLL branch) in a batch with
    two HLLC grids: every grid as on its own, bit for bit, with graph
    replay (the batch's kernel is built with Roe and runs the HLLC grids
    through HLLC)."""
    roe = dataclasses.replace(small_batch()[1], solver=setups.ROE)
    roe_prod = dataclasses.replace(setups.cvidl_wind(32, 16, dfloor=1.e-15), first_dt=1.e-3,
                                   shock_flattening=True, solver=setups.ROE)
    cfgs = [small_batch()[0], roe, small_batch()[2], roe_prod]
    V0s = [setups.wind_ic_mixed(c, seed=80 + g, rho0_cgs=1.e-11, t_range=(1.e2, 3.e7))
           for g, c in enumerate(cfgs)]
    batch = run_batch(pmx, cfgs, V0s, 6, graph=True)
    for cfg, V0, (Vb, Ub, tb) in zip(cfgs, V0s, batch):
        Vs, Us, ts = run_single(pmx, cfg, V0, 6)
        assert np.array_equal(Vb, Vs) and np.array_equal(Ub, Us)
        assert tb["dt"] == ts["dt"]
