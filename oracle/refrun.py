"""Test infrastructure: drive the CPU oracles.

* ``run_reference`` runs the unmodified PLUTO reference build
  (oracle/_ref/<problem>/pluto, compiled by oracle/ref/Makefile from the
  reference sources) on a configuration and an initial condition, and
  returns the state it dumps after the requested number of steps.
* ``run_oracle`` runs the plain-C restatement (oracle/_build/liboracle.so).

Only tests, the package smoke test and bench.py's cpu_baseline leg use
this module; the product never does.
"""
from __future__ import annotations

import ctypes
import os
import pathlib
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
REF_DIR = HERE / "_ref"
LIB_PATH = HERE / "_build" / "liboracle.so"
PKG_DIR = HERE.parent / "pluto-sirocco_amd"
if str(PKG_DIR) not in sys.path:
    sys.path.insert(0, str(PKG_DIR))


def build(with_reference: bool | None = None) -> None:
    """Compile the oracle (and, when the reference sources are present,
    the reference CPU builds)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    if with_reference is None:
        with_reference = pathlib.Path("/root/reference/Src").is_dir()
    if with_reference:
        subprocess.run(["make", "-s", "-j8", "-C", str(HERE / "ref"), "all"],
                       check=True, stdout=subprocess.DEVNULL)
        # the drop-in builds link libplutomx.so, built before this
        if (PKG_DIR / "libplutomx.so").exists():
            subprocess.run(["make", "-s", "-j8", "-C", str(HERE / "ref"), "glue"],
                           check=True, stdout=subprocess.DEVNULL)


def reference_binary(problem: str, glue: bool = False) -> pathlib.Path:
    """The reference build of a problem, or (glue) the same problem linked
    with the drop-in (pluto-sirocco_amd/glue/pmx_glue.c + libplutomx.so)."""
    return REF_DIR / (f"glue_{problem}" if glue else problem) / "pluto"


def read_dump(path: pathlib.Path, shape) -> dict:
    raw = path.read_bytes()
    magic, step, nx1, nx2, nx3, nvar = struct.unpack_from("<6i", raw, 0)
    assert magic == 0x504D5844, "bad dump magic"
    t, dt, wall = struct.unpack_from("<3d", raw, 24)
    n = nvar * nx1 * nx2 * nx3
    arr = np.frombuffer(raw, dtype="<f8", count=2 * n, offset=48)
    V = arr[:n].reshape(nvar, nx3, nx2, nx1).copy()
    U = arr[n:].reshape(nvar, nx3, nx2, nx1).copy()
    out = {"step": step, "t": t, "dt": dt, "wall": wall}
    # variables beyond the configuration's (the entropy of ENTROPY_SWITCH
    # builds) are returned apart
    nv = shape[0]
    out["V"], out["U"] = V[:nv].copy(), U[:nv].copy()
    if nvar > nv:
        out["V_extra"], out["U_extra"] = V[nv:].copy(), U[nv:].copy()
    assert out["V"].shape == tuple(shape)
    return out


def run_reference(cfg, V0: np.ndarray, nsteps: int, time_only: bool = False,
                  timeout: float = 3600.0, fluxes=None, sirocco=None,
                  t0: float | None = None, files_dir=None, output: bool = False,
                  restart=None, glue: bool = False) -> dict:
    """Run the reference for ``nsteps`` (>= 2) steps from ``V0``.

    PLUTO's ``-maxsteps N`` integrates N+1 steps (Src/main.c:233-240), so
    N = nsteps - 1 is passed.  Returns V, U (Uc), t, next dt, the dt
    sequence, and (time_only) the wall time of steps 2..nsteps.

    Wind only: ``sirocco`` (setups.SiroccoInputs) writes the files of a
    SIROCCO cycle; with heat-cool or prefactor data the run reads them as a
    restart does (restart_flag = 1).  ``t0`` starts the clock there, as a
    restart from a file at that time would.  ``files_dir``: take the
    SIROCCO files (*.dat) from there instead, as a restart.  ``output``
    (wind): the production user variables and a .dbl file every step; the
    result then holds "dbl" (the bytes of every data.NNNN.dbl, in order),
    "dbl_out" (the text of dbl.out) and "restart_out" (its bytes).
    ``restart`` = (nrestart, nmore), with output: then restart the
    reference from output nrestart of that run in the same directory
    (``-restart``) for nmore steps; "restarted" holds that run's V, U, t,
    dt and "dbl" (the files it wrote)."""
    if nsteps < 2:
        raise ValueError("the reference driver needs nsteps >= 2")
    binary = reference_binary(cfg.reference_problem(), glue)
    if not binary.exists():
        raise FileNotFoundError(f"{binary} not built (make -C oracle/ref)")
    with tempfile.TemporaryDirectory(prefix="pmxref_") as tmp:
        tmp = pathlib.Path(tmp)
        (tmp / "pluto.ini").write_text(cfg.pluto_ini(output=True) if output else cfg.pluto_ini())
        ic = np.ascontiguousarray(V0, dtype="<f8")
        if getattr(cfg, "entropy_switch", False):
            # the build carries Vc[ENTR] too; Boundary() sets it from rho, p
            ic = np.concatenate([ic, np.zeros_like(ic[:1])])
        ic.tofile(tmp / "ic.bin")
        env = dict(os.environ)
        for k in ("PMX_TIME_ONLY", "PMX_T0", "PMX_SIROCCO_RESTART"):
            env.pop(k, None)
        if files_dir is not None:
            import shutil
            for f in pathlib.Path(files_dir).glob("*.dat"):
                shutil.copy(f, tmp / f.name)
            env["PMX_SIROCCO_RESTART"] = "1"
        elif hasattr(cfg, "nangles"):  # line-driven wind: SIROCCO flux files
            import setups
            setups.write_flux_files(cfg, tmp, fluxes)
            if sirocco is not None:
                setups.write_sirocco_files(cfg, tmp, sirocco, V0)
                if sirocco.xi is not None or sirocco.pre is not None:
                    env["PMX_SIROCCO_RESTART"] = "1"
        if t0 is not None:
            env["PMX_T0"] = repr(float(t0))
        env["PMX_IC"] = str(tmp / "ic.bin")
        env["PMX_DUMP"] = str(tmp / "dump.bin")
        if time_only:
            env["PMX_TIME_ONLY"] = "1"
        proc = subprocess.run([str(binary), "-i", "pluto.ini", "-maxsteps",
                               str(nsteps - 1)], cwd=tmp, env=env,
                              capture_output=True, text=True, timeout=timeout)
        if proc.returncode != 0:
            raise RuntimeError(f"reference failed:\n{proc.stdout[-2000:]}\n{proc.stderr[-2000:]}")
        if time_only:
            walls = [float(l.split("wall=")[1]) for l in proc.stdout.splitlines()
                     if l.startswith("PMX_TIMING")]
            return {"wall": walls[-1] - walls[0], "steps_timed": len(walls) - 1}
        out = read_dump(tmp / "dump.bin", cfg.shape)
        # Analysis() runs at start-up (main.c:145-150), before steps 1..N-1
        # and after the last step: its records hold the dt of steps
        # 0..N-1 followed by the next dt.
        dts = np.loadtxt(tmp / "dump.bin.dt", ndmin=2)
        out["dt_seq"] = dts[:-1, 2]
        if output:
            out["dbl"] = [f.read_bytes() for f in sorted(tmp.glob("data.*.dbl"))]
            out["dbl_out"] = (tmp / "dbl.out").read_text()
            out["restart_out"] = (tmp / "restart.out").read_bytes()
        if output and restart is not None:
            nrestart, nmore = restart
            rec = np.frombuffer(out["restart_out"], dtype=np.dtype(
                [("nstep", "<i4"), ("nfile", "<i4", 16), ("pad", "<i4"), ("t", "<f8"),
                 ("dt", "<f8"), ("fill", "V40")]))
            s0 = int(rec["nstep"][list(rec["nfile"][:, 0]).index(nrestart)])
            before = {f.name for f in tmp.glob("data.*.dbl")}
            env["PMX_DUMP"] = str(tmp / "dump2.bin")
            proc = subprocess.run([str(binary), "-i", "pluto.ini", "-restart", str(nrestart),
                                   "-maxsteps", str(s0 + nmore - 1)], cwd=tmp, env=env,
                                  capture_output=True, text=True, timeout=timeout)
            if proc.returncode != 0:
                raise RuntimeError(f"restart failed:\n{proc.stdout[-2000:]}\n{proc.stderr[-2000:]}")
            rr = read_dump(tmp / "dump2.bin", cfg.shape)
            rr["dbl"] = [f.read_bytes() for f in sorted(tmp.glob("data.*.dbl"))]
            rr["new_dbl"] = sorted(f.name for f in tmp.glob("data.*.dbl") if f.name not in before)
            rr["step0"] = s0
            out["restarted"] = rr
        return out


# ----------------------------------------------------------------------
# plain-C restatement
# ----------------------------------------------------------------------
class _OrcCfg(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int), ("nx", ctypes.c_int * 3),
                ("recon", ctypes.c_int), ("limiter", ctypes.c_int),
                ("solver", ctypes.c_int), ("rk", ctypes.c_int),
                ("bc", ctypes.c_int * 6), ("gamma", ctypes.c_double),
                ("cfl", ctypes.c_double), ("cfl_max_var", ctypes.c_double),
                ("xbeg", ctypes.c_double * 3), ("xend", ctypes.c_double * 3),
                ("tstop", ctypes.c_double), ("first_dt", ctypes.c_double)]


# why an oracle run ended before its step count (oracle.h orc_last_quit)
QUIT_NAMES = {0: None, 1: "dt_too_small", 2: "cool_substeps", 3: "tstop"}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build(with_reference=False)
        _lib = ctypes.CDLL(str(LIB_PATH))
        dp = ctypes.POINTER(ctypes.c_double)
        _lib.orc_run.argtypes = [ctypes.POINTER(_OrcCfg), dp, ctypes.c_int, dp, dp, dp, dp]
        _lib.orc_run.restype = ctypes.c_int
        _lib.orc_riemann.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                     dp, dp, dp, dp, dp]
        _lib.orc_cons_to_prim.argtypes = [ctypes.c_double, dp, dp]
        _lib.orc_cons_to_prim.restype = ctypes.c_int
        _lib.orc_prim_to_cons.argtypes = [ctypes.c_double, dp, dp]
        _lib.orc_last_quit.restype = ctypes.c_int
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def orc_cfg(cfg) -> _OrcCfg:
    c = _OrcCfg()
    c.dims = cfg.dims
    for d in range(3):
        c.nx[d] = cfg.nx[d] if d < cfg.dims else 1
        c.xbeg[d] = cfg.xbeg[d]
        c.xend[d] = cfg.xend[d]
    c.recon, c.limiter, c.solver, c.rk = cfg.recon, cfg.limiter, cfg.solver, cfg.rk
    for s in range(6):
        c.bc[s] = cfg.bc[s]
    c.gamma, c.cfl, c.cfl_max_var = cfg.gamma, cfg.cfl, cfg.cfl_max_var
    c.tstop = getattr(cfg, "tstop", 1.e30)
    c.first_dt = cfg.first_dt
    return c


def run_oracle(cfg, V0: np.ndarray, nsteps: int, dt0: float | None = None,
               t0: float = 0.0) -> dict:
    V = np.ascontiguousarray(V0, dtype=np.float64).copy()
    U = np.empty_like(V)
    dt = np.array([cfg.first_dt if dt0 is None else dt0])
    t = np.array([t0])
    hist = np.empty(max(nsteps, 1))
    c = orc_cfg(cfg)
    rc = lib().orc_run(ctypes.byref(c), _ptr(V), nsteps, _ptr(dt), _ptr(t),
                       _ptr(U), _ptr(hist))
    if rc < 0:
        raise RuntimeError(f"orc_run failed ({rc})")
    return {"V": V, "U": U, "t": float(t[0]), "dt": float(dt[0]),
            "dt_seq": hist[:rc].copy(), "steps": rc,
            "quit": QUIT_NAMES[lib().orc_last_quit()]}


# ----------------------------------------------------------------------
# line-driven wind restatement
# ----------------------------------------------------------------------
class _OrcWindCfg(ctypes.Structure):
    _dp = ctypes.POINTER(ctypes.c_double)
    _fields_ = [("nx1", ctypes.c_int), ("nx2", ctypes.c_int), ("cooling", ctypes.c_int),
                ("bc", ctypes.c_int * 4),
                ("gamma", ctypes.c_double), ("cfl", ctypes.c_double),
                ("cfl_max_var", ctypes.c_double),
                ("mu", ctypes.c_double), ("t_iso", ctypes.c_double),
                ("l_star", ctypes.c_double), ("f_x", ctypes.c_double),
                ("f_uv", ctypes.c_double), ("t_x", ctypes.c_double),
                ("krad", ctypes.c_double), ("alpharad", ctypes.c_double),
                ("cent_mass", ctypes.c_double),
                ("unit_density", ctypes.c_double), ("unit_length", ctypes.c_double),
                ("unit_velocity", ctypes.c_double),
                ("x1", _dp), ("x1l", _dp), ("x1r", _dp), ("dx1", _dp),
                ("x2", _dp), ("x2l", _dp), ("x2r", _dp), ("dx2", _dp),
                ("nangles", ctypes.c_int),
                ("flux_r", _dp), ("flux_t", _dp), ("flux_p", _dp),
                ("dfloor", ctypes.c_double), ("shock_flattening", ctypes.c_int),
                ("char_limiting", ctypes.c_int), ("entropy_switch", ctypes.c_int),
                ("solver", ctypes.c_int),
                ("xi", _dp), ("t_r", _dp), ("pre", _dp),
                ("fm_points", ctypes.c_int), ("fm_t", _dp), ("fm_m", _dp),
                ("tstop", ctypes.c_double), ("first_dt", ctypes.c_double),
                ("rho_0", ctypes.c_double), ("rho_alpha", ctypes.c_double),
                ("disk_mdot", ctypes.c_double)]


def _wind_struct(cfg, fluxes=None, grid=None, sirocco=None):
    """The orc_wind_cfg of a WindConfig (and of the SIROCCO restart inputs
    ``sirocco``, a setups.SiroccoInputs); returns (struct, arrays to keep)."""
    import setups
    L = lib()
    if not hasattr(L, "_wind_bound"):
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_wind_run.argtypes = [ctypes.POINTER(_OrcWindCfg), dp, ctypes.c_int, dp, dp,
                                   dp, dp]
        L.orc_wind_run.restype = ctypes.c_int
        L.orc_wind_run_uv.argtypes = [ctypes.POINTER(_OrcWindCfg), dp, ctypes.c_int, dp, dp,
                                      dp, dp, dp]
        L.orc_wind_run_uv.restype = ctypes.c_int
        L.orc_wind_cool.argtypes = [ctypes.POINTER(_OrcWindCfg), dp, ctypes.c_double,
                                    ctypes.c_double, dp, dp, dp]
        L.orc_wind_cool.restype = ctypes.c_int
        L.orc_wind_cool_time.argtypes = [ctypes.POINTER(_OrcWindCfg)] + [ctypes.c_double] * 6
        L.orc_wind_cool_time.restype = ctypes.c_double
        L.orc_wind_cool_time_pre.argtypes = ([ctypes.POINTER(_OrcWindCfg)] +
                                             [ctypes.c_double] * 6 + [dp])
        L.orc_wind_cool_time_pre.restype = ctypes.c_double
        L._wind_bound = True
    gr, gt = grid if grid is not None else setups.wind_grid(cfg)
    fr, ft, fp = fluxes if fluxes is not None else setups.wind_fluxes(cfg)
    keep = []

    def arr(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return _ptr(a)

    c = _OrcWindCfg()
    c.nx1, c.nx2, c.cooling = cfg.nx[0], cfg.nx[1], cfg.cooling
    for s in range(4):
        c.bc[s] = cfg.bc[s]
    for name in ("gamma", "cfl", "cfl_max_var", "mu", "t_iso", "l_star", "f_x", "f_uv",
                 "t_x", "krad", "alpharad", "cent_mass", "unit_density", "unit_length",
                 "unit_velocity"):
        setattr(c, name, getattr(cfg, name))
    c.x1, c.x1l, c.x1r, c.dx1 = arr(gr["x"]), arr(gr["xl"]), arr(gr["xr"]), arr(gr["dx"])
    c.x2, c.x2l, c.x2r, c.dx2 = arr(gt["x"]), arr(gt["xl"]), arr(gt["xr"]), arr(gt["dx"])
    c.nangles = cfg.nangles
    c.flux_r, c.flux_t, c.flux_p = arr(fr), arr(ft), arr(fp)
    c.dfloor = cfg.dfloor
    c.shock_flattening = int(cfg.shock_flattening)
    c.char_limiting = int(cfg.char_limiting)
    c.entropy_switch = int(cfg.entropy_switch)
    c.solver = int(cfg.solver)
    c.tstop = getattr(cfg, "tstop", 1.e30)
    c.first_dt = cfg.first_dt
    c.rho_0, c.rho_alpha, c.disk_mdot = cfg.rho_0, cfg.rho_alpha, cfg.disk_mdot
    if sirocco is not None:
        if sirocco.xi is not None:
            xi, t_r = sirocco.heatcool_floored()
            c.xi, c.t_r = arr(xi), arr(t_r)
        if sirocco.pre is not None:
            pre = np.ones((6,) + tuple(np.shape(sirocco.pre)[1:]))
            pre[:len(sirocco.pre)] = sirocco.pre[:6]
            c.pre = arr(pre)
        if sirocco.fm_t is not None:
            c.fm_points = len(sirocco.fm_t)
            c.fm_t, c.fm_m = arr(sirocco.fm_t), arr(sirocco.fm_m)
    return c, keep


def run_wind_oracle(cfg, V0: np.ndarray, nsteps: int, dt0: float | None = None,
                    t0: float = 0.0, fluxes=None, grid=None, sirocco=None,
                    user_vars: bool = False) -> dict:
    """Run the wind restatement (orc_wind_run) for nsteps steps from step
    number 0.  ``fluxes`` = (flux_r, flux_t, flux_p) [nangles][1][nt][nr]
    (default: setups.wind_fluxes); ``grid`` = (radial, polar) dicts with
    x, xl, xr, dx including ghosts (default: setups.wind_grid);
    ``sirocco``: setups.SiroccoInputs of a restarted cycle; ``user_vars``:
    also ComputeUserVar on the final state ("uv", [23][nt][nr])."""
    c, keep = _wind_struct(cfg, fluxes, grid, sirocco)
    V = np.ascontiguousarray(V0, dtype=np.float64).copy()
    U = np.empty_like(V)
    dt = np.array([cfg.first_dt if dt0 is None else dt0])
    t = np.array([t0])
    hist = np.empty(max(nsteps, 1))
    uv = np.zeros((23, cfg.nx[1], cfg.nx[0])) if user_vars else None
    rc = lib().orc_wind_run_uv(ctypes.byref(c), _ptr(V), nsteps, _ptr(dt), _ptr(t), _ptr(U),
                               _ptr(hist), _ptr(uv) if uv is not None else None)
    if rc < 0:
        raise RuntimeError(f"orc_wind_run failed ({rc})")
    out = {"V": V, "U": U, "t": float(t[0]), "dt": float(dt[0]),
           "dt_seq": hist[:rc].copy(), "steps": rc,
           "quit": QUIT_NAMES[lib().orc_last_quit()]}
    if uv is not None:
        out["uv"] = uv
    return out


def run_wind_cool_oracle(cfg, V0: np.ndarray, dt: float, t: float = 0.0, xi=None, t_r=None,
                         fluxes=None, grid=None, pre=None) -> dict:
    """One SplitSource (cooling) call of the restatement on V0; ``pre``:
    the cooling prefactors [>=5][nt][nr] (default 1)."""
    import setups
    c, keep = _wind_struct(cfg, fluxes, grid,
                           None if pre is None else setups.SiroccoInputs(pre=np.asarray(pre)))
    V = np.ascontiguousarray(V0, dtype=np.float64).copy()
    dtc = np.array([0.0])
    xi_p = t_p = None
    if xi is not None:
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        t_r = np.ascontiguousarray(t_r, dtype=np.float64)
        xi_p, t_p = _ptr(xi), _ptr(t_r)
    rc = lib().orc_wind_cool(ctypes.byref(c), _ptr(V), dt, t, xi_p, t_p, _ptr(dtc))
    if rc != 0:
        raise RuntimeError(f"orc_wind_cool failed ({rc})")
    return {"V": V, "dt_cool": float(dtc[0]), "quit": QUIT_NAMES[lib().orc_last_quit()]}


WIND_COUNTERS = ("rkf12_ok", "ck45_calls", "ck45_sub", "ck45_reject", "bl_solve",
                 "bl_widen", "bl_second", "below_minT", "dt_cfl", "dt_cool", "dt_maxvar",
                 "c2p_fail", "floor_T", "floor_cells", "shock_zones", "midplane_cells")


def wind_counters() -> dict:
    """Branch counters of the last run_wind_oracle call."""
    out = (ctypes.c_long * len(WIND_COUNTERS))()
    lib().orc_wind_counters(out, len(WIND_COUNTERS))
    return dict(zip(WIND_COUNTERS, list(out)))


def wind_cool_time(cfg, rho, p0, xi, p1, tol, pre=None):
    """Exact time for p0 -> p1 under the cooling ODE (oracle
    orc_wind_cool_time); NaN if unreachable.  ``pre``: the five cooling
    prefactors (comp_h, comp_c, xray_h, brem_c, line_c), default 1."""
    c, keep = _wind_struct(cfg)
    if pre is None:
        return lib().orc_wind_cool_time(ctypes.byref(c), rho, p0, xi, cfg.t_x, p1, tol)
    pre = np.ascontiguousarray(pre, dtype=np.float64)
    return lib().orc_wind_cool_time_pre(ctypes.byref(c), rho, p0, xi, cfg.t_x, p1, tol,
                                        _ptr(pre))
