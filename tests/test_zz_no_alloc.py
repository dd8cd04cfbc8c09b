"""The hydro step without allocation (pmx_hydro_advance_step_noalloc),
called through the C ABI, against pmx_hydro_advance_step: bit for bit in
V, U and Dts->invDt_hyp, on two calls in a row on one GPU, the second on a
smaller grid than the first (the library's fixed device array still holds
the first call's data when the second one runs)."""
import ctypes
import dataclasses
import sys

import numpy as np
import pytest

import plutomx
import setups

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def cases():
    big = setups.HydroConfig(dims=3, nx=(48, 40, 36), recon=setups.PPM, solver=setups.HLLC,
                             rk=3, bc=(setups.BC_OUTFLOW,) * 6, cfl=0.3, first_dt=1.e-5)
    yield big, setups.sedov(big, seed=41), 0.25
    small = dataclasses.replace(setups.c1_kh(32), nx=(40, 24, 1), first_dt=1.e-3)
    yield small, setups.kelvin_helmholtz(small, seed=42), 0.0


def with_context(cfg, V0, inv_in):
    with plutomx.Hydro(cfg) as h:
        h.set_prim(V0)
        h.set_time(0.0, cfg.first_dt)
        inv = ctypes.c_double(inv_in)
        lib = plutomx.load_library()
        assert lib.pmx_hydro_advance_step(h._h, cfg.first_dt, ctypes.byref(inv)) == 0
        return h.get_prim(), h.get_cons(), inv.value


def without_allocation(cfg, V0, inv_in):
    lib = plutomx.load_library()
    keep = []
    d = plutomx.hydro_desc(cfg, keep)
    with plutomx.Hydro(cfg) as h:    # its device buffers serve as the caller's V and U
        h.set_prim(V0)
        V, U = h.device_state()
        inv = ctypes.c_double(inv_in)
        rc = lib.pmx_hydro_advance_step_noalloc(ctypes.byref(d), V, U, cfg.first_dt,
                                                ctypes.byref(inv))
        assert rc == 0, lib.pmx_last_error()
        return h.get_prim(), h.get_cons(), inv.value


def test_noalloc_matches_advance_step_on_two_calls():
    for n, (cfg, V0, inv_in) in enumerate(cases()):
        say(f"[gpu] no-allocation step, call {n + 1}: {cfg.nx}, dims {cfg.dims}")
        Vb, Ub, ib = without_allocation(cfg, V0, inv_in)
        Va, Ua, ia = with_context(cfg, V0, inv_in)
        assert np.isfinite(Va).all() and np.isfinite(Ua).all()
        assert np.array_equal(Va, Vb), np.abs(Va - Vb).max()
        assert np.array_equal(Ua, Ub), np.abs(Ua - Ub).max()
        assert ia == ib and ia > 0.0
