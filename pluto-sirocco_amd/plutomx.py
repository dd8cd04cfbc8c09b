"""Host-side Python binding of libplutomx (ctypes over include/plutomx.h).

This mirrors the reference's driver interface for the hot path:
``Hydro.advance_step(dt)`` is PLUTO's ``AdvanceStep`` (returning
``Dts->invDt_hyp``), ``Hydro.run(n)`` is ``n`` iterations of the main loop
(``Integrate`` + ``NextTimeStep``) kept on the device.  There is no CPU
fallback: if the HIP library is missing or no GPU is visible, every call
raises.
"""
from __future__ import annotations

import ctypes
import os
import pathlib
import weakref

import numpy as np

import setups

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libplutomx.so"

REC_CODE = {  # (recon, limiter) -> PMX_REC_*
    (setups.PLM, setups.LIM_VANLEER): 0,
    (setups.PLM, setups.LIM_MC): 1,
    (setups.PLM, setups.LIM_MINMOD): 2,
    (setups.PLM, setups.LIM_DEFAULT): 3,
}
PMX_REC_PPM = 5

EXPORTS = ["pmx_last_error", "pmx_hydro_create", "pmx_hydro_destroy", "pmx_hydro_set_prim",
           "pmx_hydro_get_prim", "pmx_hydro_get_cons", "pmx_hydro_set_time",
           "pmx_hydro_get_time", "pmx_hydro_advance_step", "pmx_hydro_run",
           "pmx_hydro_sync", "pmx_hydro_run_graph", "pmx_hydro_stage_timing",
           "pmx_hydro_device_state", "pmx_copy_probe", "pmx_copy_probe_runtime", "pmx_check_div_rc",
           "pmx_check_force_power",
           "pmx_wind_create", "pmx_wind_destroy", "pmx_wind_set_prim", "pmx_wind_get_prim",
           "pmx_wind_get_cons", "pmx_wind_set_heatcool", "pmx_wind_set_time",
           "pmx_wind_get_time", "pmx_wind_set_dts", "pmx_wind_advance_step", "pmx_wind_split_source",
           "pmx_wind_run", "pmx_wind_sync", "pmx_wind_timing",
           "pmx_mlp_create", "pmx_mlp_destroy", "pmx_mlp_set_params", "pmx_mlp_get_params",
           "pmx_mlp_sample", "pmx_mlp_set_dataset", "pmx_mlp_get_dataset",
           "pmx_mlp_loss_grad", "pmx_mlp_train", "pmx_mlp_predict",
           "pmx_wind_set_surrogate", "pmx_wind_surrogate_stats", "pmx_mlp_cool_cells",
           "pmx_wind_batch_create", "pmx_wind_batch_destroy", "pmx_wind_batch_grid",
           "pmx_wind_batch_set_step", "pmx_wind_batch_set_surrogate", "pmx_wind_batch_autotune",
           "pmx_wind_batch_set_shape", "pmx_wind_batch_run", "pmx_wind_batch_sync",
           "pmx_wind_batch_timing", "pmx_wind_batch_cells",
           "pmx_wind_set_prefactors", "pmx_sirocco_flux_angles", "pmx_sirocco_read_fluxes",
           "pmx_sirocco_fm_points", "pmx_sirocco_read_force_multiplier",
           "pmx_sirocco_read_heatcool", "pmx_sirocco_read_prefactors",
           "pmx_wind_snapshot", "pmx_wind_snapshot_wait", "pmx_write_dbl", "pmx_write_restart",
           "pmx_read_restart", "pmx_read_dbl", "pmx_wind_snapshot_host", "pmx_wind_run_graph",
           "pmx_hydro_set_tstop", "pmx_wind_set_tstop", "pmx_hydro_advance_step_noalloc"]


class PmxError(RuntimeError):
    pass


# Run-stopping flags (sticky; the main loop's steps are no-ops once one is
# set, until set_time), as in include/plutomx.h.
PMX_FLAG_COOL_SUBSTEPS = 256   # a Cash-Karp cooling integration exceeded 1000 sub-steps
PMX_FLAG_DT_TOO_SMALL = 4096   # NextTimeStep stopped: dt < 1e-9 first_dt
PMX_FLAG_TSTOP = 8192          # the step that reached tstop has been taken
PMX_FLAGS_HALT = PMX_FLAG_COOL_SUBSTEPS | PMX_FLAG_DT_TOO_SMALL | PMX_FLAG_TSTOP


class HydroDesc(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int), ("nx1", ctypes.c_int), ("nx2", ctypes.c_int),
                ("nx3", ctypes.c_int), ("recon", ctypes.c_int), ("solver", ctypes.c_int),
                ("rk", ctypes.c_int), ("bc", ctypes.c_int * 6), ("gamma", ctypes.c_double),
                ("cfl", ctypes.c_double), ("cfl_max_var", ctypes.c_double),
                ("dx1", ctypes.POINTER(ctypes.c_double)),
                ("dx2", ctypes.POINTER(ctypes.c_double)),
                ("dx3", ctypes.POINTER(ctypes.c_double)),
                ("first_dt", ctypes.c_double)]


class WindDesc(ctypes.Structure):
    _dp = ctypes.POINTER(ctypes.c_double)
    _fields_ = [("nx1", ctypes.c_int), ("nx2", ctypes.c_int), ("cooling", ctypes.c_int),
                ("bc", ctypes.c_int * 4),
                ("gamma", ctypes.c_double), ("cfl", ctypes.c_double),
                ("cfl_max_var", ctypes.c_double),
                ("mu", ctypes.c_double), ("l_star", ctypes.c_double), ("f_xray", ctypes.c_double),
                ("t_x", ctypes.c_double), ("krad", ctypes.c_double),
                ("alpharad", ctypes.c_double), ("cent_mass", ctypes.c_double),
                ("unit_density", ctypes.c_double), ("unit_length", ctypes.c_double),
                ("unit_velocity", ctypes.c_double),
                ("x1", _dp), ("x1l", _dp), ("x1r", _dp), ("dx1", _dp),
                ("x2", _dp), ("x2l", _dp), ("x2r", _dp), ("dx2", _dp),
                ("nangles", ctypes.c_int),
                ("flux_r", _dp), ("flux_t", _dp), ("flux_p", _dp),
                ("dfloor", ctypes.c_double), ("shock_flattening", ctypes.c_int),
                ("char_limiting", ctypes.c_int), ("entropy_switch", ctypes.c_int),
                ("solver", ctypes.c_int), ("first_dt", ctypes.c_double),
                ("fm_points", ctypes.c_int), ("fm_t", _dp), ("fm_m", _dp),
                ("force_power", ctypes.c_int),
                ("rho_0", ctypes.c_double), ("rho_alpha", ctypes.c_double),
                ("disk_mdot", ctypes.c_double)]

PMX_FORCE_POWER_REF = 0        # the reference's t and pow(t, alpha) per LineForce call
PMX_FORCE_POWER_FACTORED = 1   # one exp2(-alpha log2 dvds) per bin for both calls


_lib = None


def load_library(path: pathlib.Path | str | None = None) -> ctypes.CDLL:
    """Load libplutomx.so (built by build.py); raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = pathlib.Path(path) if path else pathlib.Path(os.environ.get("PMX_LIB", LIB_PATH))
    if not p.exists():
        raise PmxError(f"{p} is not built: run `python pluto-sirocco_amd/build.py`")
    lib = ctypes.CDLL(str(p))
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    vp = ctypes.c_void_p
    lib.pmx_last_error.restype = ctypes.c_char_p
    lib.pmx_hydro_create.argtypes = [ctypes.POINTER(HydroDesc), ctypes.POINTER(vp)]
    lib.pmx_hydro_destroy.argtypes = [vp]
    lib.pmx_hydro_set_prim.argtypes = [vp, dp, ctypes.c_int]
    lib.pmx_hydro_get_prim.argtypes = [vp, dp, ctypes.c_int]
    lib.pmx_hydro_get_cons.argtypes = [vp, dp, ctypes.c_int]
    lib.pmx_hydro_set_time.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    lib.pmx_hydro_set_tstop.argtypes = [vp, ctypes.c_double]
    lib.pmx_hydro_advance_step_noalloc.argtypes = [ctypes.POINTER(HydroDesc), vp, vp,
                                                   ctypes.c_double, dp]
    lib.pmx_hydro_get_time.argtypes = [vp, dp, dp, dp, ip]
    lib.pmx_hydro_advance_step.argtypes = [vp, ctypes.c_double, dp]
    lib.pmx_hydro_run.argtypes = [vp, ctypes.c_int]
    lib.pmx_hydro_run_graph.argtypes = [vp, ctypes.c_int]
    lib.pmx_hydro_sync.argtypes = [vp]
    lib.pmx_hydro_stage_timing.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, ip]
    lib.pmx_hydro_device_state.argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp)]
    lib.pmx_copy_probe.argtypes = [ctypes.c_longlong, ctypes.c_int, dp]
    lib.pmx_copy_probe_runtime.argtypes = [ctypes.c_longlong, ctypes.c_int, dp]
    lib.pmx_check_div_rc.argtypes = [dp, ctypes.c_longlong, ctypes.c_double,
                                      ctypes.POINTER(ctypes.c_longlong)]
    lib.pmx_check_force_power.argtypes = [dp, dp, ctypes.c_longlong, ctypes.c_double,
                                          ctypes.c_double, dp]
    lp = ctypes.POINTER(ctypes.c_long)
    lib.pmx_wind_create.argtypes = [ctypes.POINTER(WindDesc), ctypes.POINTER(vp)]
    lib.pmx_wind_destroy.argtypes = [vp]
    lib.pmx_wind_set_prim.argtypes = [vp, dp]
    lib.pmx_wind_get_prim.argtypes = [vp, dp]
    lib.pmx_wind_get_cons.argtypes = [vp, dp]
    lib.pmx_wind_set_heatcool.argtypes = [vp, dp, dp]
    lib.pmx_wind_set_time.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_long]
    lib.pmx_wind_set_tstop.argtypes = [vp, ctypes.c_double]
    lib.pmx_wind_get_time.argtypes = [vp, dp, dp, lp, dp, dp, ip]
    lib.pmx_wind_set_dts.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    lib.pmx_wind_advance_step.argtypes = [vp]
    lib.pmx_wind_split_source.argtypes = [vp]
    lib.pmx_wind_run.argtypes = [vp, ctypes.c_int]
    lib.pmx_wind_run_graph.argtypes = [vp, ctypes.c_int]
    lib.pmx_wind_sync.argtypes = [vp]
    lib.pmx_wind_timing.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, ip, dp, ip]
    fp = ctypes.POINTER(ctypes.c_float)
    ll = ctypes.c_longlong
    lib.pmx_mlp_create.argtypes = [ctypes.c_ulonglong, ctypes.POINTER(vp)]
    lib.pmx_mlp_destroy.argtypes = [vp]
    lib.pmx_mlp_set_params.argtypes = [vp, fp]
    lib.pmx_mlp_get_params.argtypes = [vp, fp]
    lib.pmx_mlp_sample.argtypes = [vp, ctypes.POINTER(WindDesc), ll, ctypes.c_ulonglong]
    lib.pmx_mlp_set_dataset.argtypes = [vp, fp, fp, ll]
    lib.pmx_mlp_get_dataset.argtypes = [vp, fp, fp, ip, dp, ll]
    lib.pmx_mlp_loss_grad.argtypes = [vp, ll, fp, fp]
    lib.pmx_mlp_train.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                  ctypes.c_float, fp]
    lib.pmx_mlp_predict.argtypes = [vp, fp, ll, fp]
    llp = ctypes.POINTER(ctypes.c_longlong)
    lib.pmx_wind_set_surrogate.argtypes = [vp, vp, ctypes.c_int]
    lib.pmx_wind_batch_create.argtypes = [ctypes.POINTER(WindDesc), ctypes.c_int,
                                          ctypes.POINTER(vp)]
    lib.pmx_wind_batch_destroy.argtypes = [vp]
    lib.pmx_wind_batch_grid.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp)]
    lib.pmx_wind_batch_set_step.argtypes = [vp, ctypes.c_long]
    lib.pmx_wind_batch_set_surrogate.argtypes = [vp, vp, ctypes.c_int]
    lib.pmx_wind_batch_autotune.argtypes = [vp, ip, ip, dp]
    lib.pmx_wind_batch_set_shape.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    lib.pmx_wind_batch_run.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    lib.pmx_wind_batch_sync.argtypes = [vp]
    lib.pmx_wind_batch_timing.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, ip, dp, ip]
    lib.pmx_wind_batch_cells.argtypes = [vp, llp]
    lib.pmx_mlp_cool_cells.argtypes = [vp, ctypes.POINTER(WindDesc), ll, dp, dp, dp, dp, dp,
                                       ctypes.c_int, llp, dp]
    lib.pmx_wind_surrogate_stats.argtypes = [vp, llp, llp, llp, llp, ctypes.c_int]
    lib.pmx_wind_set_prefactors.argtypes = [vp, dp]
    cp = ctypes.c_char_p
    wdp = ctypes.POINTER(WindDesc)
    lib.pmx_sirocco_flux_angles.argtypes = [cp, ip]
    lib.pmx_sirocco_read_fluxes.argtypes = [cp, wdp, ctypes.c_int, dp, dp, dp, ip, lp]
    lib.pmx_sirocco_fm_points.argtypes = [cp, ip]
    lib.pmx_sirocco_read_force_multiplier.argtypes = [cp, wdp, ctypes.c_int, dp, dp, ip, lp]
    lib.pmx_sirocco_read_heatcool.argtypes = [cp, wdp, dp, dp, dp, ip, lp]
    lib.pmx_sirocco_read_prefactors.argtypes = [cp, wdp, dp, ip, lp, lp]
    lib.pmx_wind_snapshot.argtypes = [vp, ctypes.c_int]
    lib.pmx_wind_snapshot_wait.argtypes = [vp, dp, dp, dp, lp]
    lib.pmx_wind_snapshot_host.argtypes = [vp, ctypes.POINTER(dp)]
    lib.pmx_write_dbl.argtypes = [cp, ctypes.c_int, dp, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_char_p), ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_long]
    lib.pmx_write_restart.argtypes = [cp, ctypes.c_int, ctypes.c_long, ip, ctypes.c_double,
                                      ctypes.c_double]
    lib.pmx_read_restart.argtypes = [cp, ctypes.c_int, lp, ip, dp, dp, ip]
    lib.pmx_read_dbl.argtypes = [cp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, dp]
    for name in EXPORTS:
        if name != "pmx_last_error":
            getattr(lib, name).restype = ctypes.c_int
    if path is None:
        _lib = lib
    return lib


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise PmxError(f"{what}: {load_library().pmx_last_error().decode()}")


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def cell_widths(cfg: setups.HydroConfig, d: int) -> np.ndarray:
    """grid->dx[d][IBEG..IEND] of a uniform grid (Src/set_grid.c:396-403)."""
    n = cfg.nx[d] if d < cfg.dims else 1
    if d >= cfg.dims:
        return np.ones(1)
    return np.full(n, (cfg.xend[d] - cfg.xbeg[d]) / float(n))


def check_div_rc(x: np.ndarray, c: float) -> int:
    """Number of x whose device quotient x / c from the reciprocal-and-FMA
    evaluation differs in any bit from the device division."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    bad = ctypes.c_longlong()
    _check(load_library().pmx_check_div_rc(_dptr(x), x.size, float(c), ctypes.byref(bad)),
           "check_div_rc")
    return bad.value


def check_force_power(num: np.ndarray, dvds: np.ndarray, krad: float, alpha: float) -> np.ndarray:
    """The device's line-force power krad pow(num / dvds, alpha) per element
    (0 where dvds <= 0), as the line-force kernel evaluates it."""
    num = np.ascontiguousarray(num, dtype=np.float64)
    dvds = np.ascontiguousarray(dvds, dtype=np.float64)
    if num.shape != dvds.shape:
        raise ValueError("num and dvds differ in shape")
    out = np.empty_like(num)
    _check(load_library().pmx_check_force_power(_dptr(num), _dptr(dvds), num.size, float(krad),
                                                float(alpha), _dptr(out)), "check_force_power")
    return out


def copy_probe(nbytes: int, reps: int = 10, runtime: bool = False) -> float:
    """Mean ms of a plain device copy of nbytes (bandwidth reference): the
    library's streaming copy kernel, or (runtime) the HIP runtime's own
    device-to-device hipMemcpyAsync."""
    ms = ctypes.c_double()
    fn = load_library().pmx_copy_probe_runtime if runtime else load_library().pmx_copy_probe
    _check(fn(nbytes, reps, ctypes.byref(ms)), "copy_probe")
    return ms.value


def hydro_desc(cfg: setups.HydroConfig, keep: list) -> HydroDesc:
    """The pmx_hydro_desc of a configuration; the cell-width arrays it
    points to are appended to `keep` (they must outlive its use)."""
    d = HydroDesc()
    d.dims = cfg.dims
    d.nx1, d.nx2 = cfg.nx[0], cfg.nx[1]
    d.nx3 = cfg.nx[2] if cfg.dims == 3 else 1
    d.recon = PMX_REC_PPM if cfg.recon == setups.PPM else REC_CODE[(cfg.recon, cfg.limiter)]
    d.solver, d.rk = cfg.solver, cfg.rk
    for s in range(6):
        d.bc[s] = cfg.bc[s]
    d.gamma, d.cfl, d.cfl_max_var = cfg.gamma, cfg.cfl, cfg.cfl_max_var
    d.first_dt = cfg.first_dt
    dx = [cell_widths(cfg, q) for q in range(3)]
    keep.extend(dx)
    d.dx1, d.dx2, d.dx3 = (_dptr(x) for x in dx)
    return d


class Hydro:
    """One grid advanced on the GPU by libplutomx."""

    def __init__(self, cfg: setups.HydroConfig):
        self.cfg = cfg
        lib = load_library()
        self._dx = []
        d = hydro_desc(cfg, self._dx)
        h = ctypes.c_void_p()
        _check(lib.pmx_hydro_create(ctypes.byref(d), ctypes.byref(h)), "pmx_hydro_create")
        self._h = h
        self._lib = lib
        self.set_tstop(getattr(cfg, "tstop", 1.e30))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.pmx_hydro_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- state -----------------------------------------------------------
    def set_prim(self, V: np.ndarray, ghosts: bool = False) -> None:
        V = np.ascontiguousarray(V, dtype=np.float64)
        if V.size != self._prim_size(ghosts):
            raise PmxError(f"set_prim: {V.size} values, the grid holds {self._prim_size(ghosts)}")
        _check(self._lib.pmx_hydro_set_prim(self._h, _dptr(V), int(ghosts)), "set_prim")

    def _prim_size(self, ghosts: bool) -> int:
        nv, nz, ny, nx = self.cfg.shape
        if not ghosts:
            return nv * nz * ny * nx
        g = self.cfg.nghost
        return nv * (nz + (2 * g if self.cfg.dims == 3 else 0)) * (ny + 2 * g) * (nx + 2 * g)

    def get_prim(self, ghosts: bool = False) -> np.ndarray:
        shape = self.cfg.shape
        if ghosts:
            g = self.cfg.nghost
            shape = (5, shape[1] + (2 * g if self.cfg.dims == 3 else 0), shape[2] + 2 * g,
                     shape[3] + 2 * g)
        out = np.empty(shape)
        _check(self._lib.pmx_hydro_get_prim(self._h, _dptr(out), int(ghosts)), "get_prim")
        return out

    def get_cons(self) -> np.ndarray:
        out = np.empty(self.cfg.shape)
        _check(self._lib.pmx_hydro_get_cons(self._h, _dptr(out), 0), "get_cons")
        return out

    def set_time(self, t: float, dt: float) -> None:
        _check(self._lib.pmx_hydro_set_time(self._h, t, dt), "set_time")

    def set_tstop(self, tstop: float) -> None:
        """runtime.tstop: run() cuts the step that reaches it and halts
        after it (PMX_FLAG_TSTOP)."""
        _check(self._lib.pmx_hydro_set_tstop(self._h, tstop), "set_tstop")

    def get_time(self):
        t, dt, inv = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        fl = ctypes.c_int()
        _check(self._lib.pmx_hydro_get_time(self._h, ctypes.byref(t), ctypes.byref(dt),
                                            ctypes.byref(inv), ctypes.byref(fl)), "get_time")
        return t.value, dt.value, inv.value, fl.value

    # --- stepping --------------------------------------------------------
    def advance_step(self, dt: float) -> float:
        inv = ctypes.c_double()
        _check(self._lib.pmx_hydro_advance_step(self._h, dt, ctypes.byref(inv)), "advance_step")
        return inv.value

    def run(self, nsteps: int, graph: bool = False) -> None:
        fn = self._lib.pmx_hydro_run_graph if graph else self._lib.pmx_hydro_run
        _check(fn(self._h, nsteps), "run")

    def sync(self) -> None:
        _check(self._lib.pmx_hydro_sync(self._h), "sync")

    def device_state(self):
        """Device addresses of the current V and U ([nv][k][j][i])."""
        V, U = ctypes.c_void_p(), ctypes.c_void_p()
        _check(self._lib.pmx_hydro_device_state(self._h, ctypes.byref(V), ctypes.byref(U)),
               "device_state")
        return V, U

    def stage_timing(self, enable: bool, reset: bool = True):
        ms, n = ctypes.c_double(), ctypes.c_int()
        _check(self._lib.pmx_hydro_stage_timing(self._h, int(enable), int(reset),
                                                ctypes.byref(ms), ctypes.byref(n)), "timing")
        return ms.value, n.value


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def wind_desc_scalars(cfg: setups.WindConfig) -> WindDesc:
    """A descriptor carrying only the scalar physics of a wind setup (units,
    gamma, mu, T_x, ...), as the surrogate's sampler needs."""
    d = WindDesc()
    d.nx1, d.nx2, d.cooling = cfg.nx[0], cfg.nx[1], cfg.cooling
    for q in range(4):
        d.bc[q] = cfg.bc[q]
    for name in ("gamma", "cfl", "cfl_max_var", "mu", "l_star", "t_x", "krad",
                 "alpharad", "cent_mass", "unit_density", "unit_length", "unit_velocity"):
        setattr(d, name, getattr(cfg, name))
    d.f_xray = cfg.f_x
    d.nangles = cfg.nangles
    d.dfloor = cfg.dfloor
    d.shock_flattening = int(cfg.shock_flattening)
    d.char_limiting = int(cfg.char_limiting)
    d.entropy_switch = int(cfg.entropy_switch)
    d.solver = int(cfg.solver)
    d.first_dt = cfg.first_dt
    d.force_power = int(getattr(cfg, "force_power", PMX_FORCE_POWER_REF))
    d.rho_0, d.rho_alpha, d.disk_mdot = cfg.rho_0, cfg.rho_alpha, cfg.disk_mdot
    return d


class Mlp:
    """The MLP cooling surrogate (4 -> 64 -> 64 -> 64 -> 1, SiLU, BF16
    matrix-core operands, FP32 accumulation) trained on the GPU."""

    NP = 8705

    def __init__(self, seed: int = 1):
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib.pmx_mlp_create(seed, ctypes.byref(h)), "pmx_mlp_create")
        self._h, self._lib = h, lib

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.pmx_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def get_params(self) -> np.ndarray:
        P = np.empty(self.NP, dtype=np.float32)
        _check(self._lib.pmx_mlp_get_params(self._h, _fptr(P)), "get_params")
        return P

    def set_params(self, P: np.ndarray) -> None:
        P = np.ascontiguousarray(P, dtype=np.float32)
        _check(self._lib.pmx_mlp_set_params(self._h, _fptr(P)), "set_params")

    def sample(self, cfg: setups.WindConfig, n: int, seed: int = 7) -> None:
        d = wind_desc_scalars(cfg)
        _check(self._lib.pmx_mlp_sample(self._h, ctypes.byref(d), n, seed), "sample")

    def set_dataset(self, X: np.ndarray, Y: np.ndarray) -> None:
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.ascontiguousarray(Y, dtype=np.float32)
        _check(self._lib.pmx_mlp_set_dataset(self._h, _fptr(X), _fptr(Y), len(Y)), "set_dataset")

    def get_dataset(self, n: int):
        """(X[n][4], Y[n], nev[n], aux[n][4]) of the first n samples."""
        X = np.empty((n, 4), dtype=np.float32)
        Y = np.empty(n, dtype=np.float32)
        nev = np.empty(n, dtype=np.int32)
        aux = np.empty((n, 4), dtype=np.float64)
        _check(self._lib.pmx_mlp_get_dataset(self._h, _fptr(X), _fptr(Y),
                                             nev.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                             _dptr(aux), n), "get_dataset")
        return X, Y, nev, aux

    def loss_grad(self, count: int):
        g = np.empty(self.NP, dtype=np.float32)
        loss = ctypes.c_float()
        _check(self._lib.pmx_mlp_loss_grad(self._h, count, ctypes.byref(loss), _fptr(g)),
               "loss_grad")
        return loss.value, g

    def train(self, steps: int, batch: int = 4096, lr0: float = 3e-3,
              lr1: float = 1e-5) -> np.ndarray:
        hist = np.empty(steps, dtype=np.float32)
        _check(self._lib.pmx_mlp_train(self._h, steps, batch, lr0, lr1, _fptr(hist)), "train")
        return hist

    def cool_cells(self, cfg: setups.WindConfig, rho, p, xi, dt, mode: int = 1):
        """One cooling step of independent cells (code units): mode 0 the
        plain integrator, 1 the surrogate, 2 the surrogate without the
        network.  Returns (p_new, stats dict, device ms)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, p, xi, dt)]
        n = len(arrs[0])
        out = np.empty(n)
        st = (ctypes.c_longlong * 4)()
        ms = ctypes.c_double()
        d = wind_desc_scalars(cfg)
        _check(self._lib.pmx_mlp_cool_cells(self._h, ctypes.byref(d), n, *(_dptr(a) for a in arrs),
                                            _dptr(out), mode, st, ctypes.byref(ms)), "cool_cells")
        return out, {"listed": st[0], "accepted": st[1], "fallback": st[2],
                     "radiat_calls": st[3]}, ms.value

    def predict(self, X: np.ndarray) -> np.ndarray:
        """log10(T_final / T_0) predicted for raw features X[n][4]."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.empty(len(X), dtype=np.float32)
        _check(self._lib.pmx_mlp_predict(self._h, _fptr(X), len(X), _fptr(Y)), "predict")
        return Y


def wind_desc(cfg: setups.WindConfig, fluxes=None, grid=None, keep=None,
              sirocco=None) -> WindDesc:
    """Full descriptor of a wind grid; the arrays it points to are appended
    to `keep` (they must outlive the create call).  ``sirocco``
    (setups.SiroccoInputs) supplies the force multiplier fits of a grid
    with krad = alpharad = 999."""
    gr, gt = grid if grid is not None else setups.wind_grid(cfg)
    fr, ft, fp = fluxes if fluxes is not None else setups.wind_fluxes(cfg)
    keep = [] if keep is None else keep

    def arr(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return _dptr(a)

    d = wind_desc_scalars(cfg)
    d.x1, d.x1l, d.x1r, d.dx1 = arr(gr["x"]), arr(gr["xl"]), arr(gr["xr"]), arr(gr["dx"])
    d.x2, d.x2l, d.x2r, d.dx2 = arr(gt["x"]), arr(gt["xl"]), arr(gt["xr"]), arr(gt["dx"])
    d.flux_r, d.flux_t, d.flux_p = arr(fr), arr(ft), arr(fp)
    if sirocco is not None and sirocco.fm_t is not None:
        d.fm_points = len(sirocco.fm_t)
        d.fm_t, d.fm_m = arr(sirocco.fm_t), arr(sirocco.fm_m)
    return d


PRIM_NAMES = ("rho", "vx1", "vx2", "vx3", "prs", "tr1")   # Src/var_names.c:33-95


def write_dbl(directory, nfile: int, data: np.ndarray, names, t: float, dt: float,
              step: int) -> None:
    """data.NNNN.dbl (data[nvar][nz][ny][nx], single_file) and its dbl.out
    line, as the reference's WriteData writes them (pmx_write_dbl)."""
    data = np.ascontiguousarray(data, dtype=np.float64)
    nvar = len(names)
    if data.shape[0] != nvar:
        raise PmxError("one name per variable")
    shape = data.shape[1:]
    nz, ny, nx = (1,) * (3 - len(shape)) + tuple(shape)
    cn = (ctypes.c_char_p * nvar)(*[n.encode() for n in names])
    _check(load_library().pmx_write_dbl(str(directory).encode(), nfile, _dptr(data), nvar, cn,
                                        nx, ny, nz, t, dt, step), "write_dbl")


def write_restart(directory, counter: int, nstep: int, nfile, t: float, dt: float) -> None:
    """Record `counter` of restart.out (pmx_write_restart); nfile: 16 file
    counters, one per output type."""
    nf = (ctypes.c_int * 16)(*[int(v) for v in nfile])
    _check(load_library().pmx_write_restart(str(directory).encode(), counter, nstep, nf, t, dt),
           "write_restart")


def read_restart(directory, nrestart: int, nvar: int, shape) -> dict:
    """Restart from output nrestart as the reference does (pmx_read_restart,
    pmx_read_dbl): its step, file counters, t, dt and the first nvar arrays
    of its .dbl file, data[nvar][*shape]."""
    lib = load_library()
    st, t, dt, sw = ctypes.c_long(), ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    nf = (ctypes.c_int * 16)()
    d = str(directory).encode()
    _check(lib.pmx_read_restart(d, nrestart, ctypes.byref(st), nf, ctypes.byref(t),
                                ctypes.byref(dt), ctypes.byref(sw)), "read_restart")
    shape = tuple(shape)
    nz, ny, nx = (1,) * (3 - len(shape)) + shape
    data = np.empty((nvar,) + shape)
    _check(lib.pmx_read_dbl(d, nf[0], nvar, nx, ny, nz, sw.value, _dptr(data)), "read_dbl")
    return {"nstep": st.value, "nfile": list(nf), "t": t.value, "dt": dt.value,
            "swap": sw.value, "data": data}


def _grid_desc(cfg: setups.WindConfig, keep: list) -> WindDesc:
    """Descriptor with the scalars and the cell centres of a wind grid (what
    the SIROCCO file readers use)."""
    gr, gt = setups.wind_grid(cfg)
    d = wind_desc_scalars(cfg)
    for a in (np.ascontiguousarray(gr["x"]), np.ascontiguousarray(gt["x"])):
        keep.append(a)
    d.x1, d.x2 = _dptr(keep[-2]), _dptr(keep[-1])
    return d


def read_sirocco_files(cfg: setups.WindConfig, directory, V: np.ndarray | None = None) -> dict:
    """Read the files of a SIROCCO cycle in ``directory`` with the
    library's readers (read_sirocco_fluxes and read_sirocco_heatcool with
    restart_flag = 1).  Returns "fluxes" (flux_r, flux_t, flux_p)
    [nangles][1][nt][nr] or None, "inputs" (setups.SiroccoInputs: xi, t_r
    after the reader's floors, pre [6][nt][nr] (all 1 without the file),
    fm_t, fm_m), and per file "found" / "matched" records.  V (the state,
    for the start-up xi / T_r when py_heatcool.dat is missing) defaults to
    no heat-cool arrays when py_heatcool.dat is missing."""
    lib = load_library()
    keep: list = []
    g = _grid_desc(cfg, keep)
    dirb = str(directory).encode()
    nt, nr = cfg.nx[1], cfg.nx[0]
    out = {"found": {}, "matched": {}}
    na = ctypes.c_int()
    _check(lib.pmx_sirocco_flux_angles(dirb, ctypes.byref(na)), "flux_angles")
    out["fluxes"] = None
    if na.value:
        fl = [np.zeros((na.value, 1, nt, nr)) for _ in range(3)]
        fnd, m = ctypes.c_int(), (ctypes.c_long * 3)()
        _check(lib.pmx_sirocco_read_fluxes(dirb, ctypes.byref(g), na.value, *map(_dptr, fl),
                                           ctypes.byref(fnd), m), "read_fluxes")
        out["fluxes"] = tuple(fl)
        out["found"]["fluxes"], out["matched"]["fluxes"] = fnd.value, list(m)
    inp = setups.SiroccoInputs()
    npts = ctypes.c_int()
    _check(lib.pmx_sirocco_fm_points(dirb, ctypes.byref(npts)), "fm_points")
    if npts.value:
        t = np.zeros(npts.value)
        m_ = np.zeros((npts.value, nt, nr))
        fnd, cnt = ctypes.c_int(), ctypes.c_long()
        _check(lib.pmx_sirocco_read_force_multiplier(dirb, ctypes.byref(g), npts.value, _dptr(t),
                                                     _dptr(m_), ctypes.byref(fnd),
                                                     ctypes.byref(cnt)), "read_force_multiplier")
        inp.fm_t, inp.fm_m = t, m_
        out["found"]["fm"], out["matched"]["fm"] = fnd.value, cnt.value
    if V is not None or (pathlib.Path(directory) / "py_heatcool.dat").exists():
        xi, t_r = np.zeros((nt, nr)), np.zeros((nt, nr))
        rho = None if V is None else np.ascontiguousarray(V[setups.RHO, 0], dtype=np.float64)
        fnd, cnt = ctypes.c_int(), ctypes.c_long()
        _check(lib.pmx_sirocco_read_heatcool(dirb, ctypes.byref(g),
                                             None if rho is None else _dptr(rho), _dptr(xi),
                                             _dptr(t_r), ctypes.byref(fnd), ctypes.byref(cnt)),
               "read_heatcool")
        inp.xi, inp.t_r = xi, t_r
        out["found"]["heatcool"], out["matched"]["heatcool"] = fnd.value, cnt.value
    pre = np.zeros((6, nt, nr))
    fnd, cnt, neg = ctypes.c_int(), ctypes.c_long(), ctypes.c_long()
    _check(lib.pmx_sirocco_read_prefactors(dirb, ctypes.byref(g), _dptr(pre), ctypes.byref(fnd),
                                           ctypes.byref(cnt), ctypes.byref(neg)), "read_prefactors")
    inp.pre = pre
    out["found"]["prefactors"], out["matched"]["prefactors"] = fnd.value, cnt.value
    out["negative_prefactors"] = neg.value
    out["inputs"] = inp
    return out


class _PinnedView:
    """Owner (numpy base) of a zero-copy, read-only view of a context's
    pinned snapshot buffer; the context tracks these through weak
    references."""

    def __init__(self, ptr, shape):
        self.__array_interface__ = {"data": (ctypes.cast(ptr, ctypes.c_void_p).value, True),
                                    "shape": shape, "typestr": "<f8", "version": 3}


class _DeferredDestroy:
    """Destroys a wind context once every listed snapshot view is gone."""

    def __init__(self, lib, handle, views):
        self._lib, self._h, self._left = lib, handle, len(views)
        for v in views:
            weakref.finalize(v, self._dropped)

    def _dropped(self):
        self._left -= 1
        if self._left == 0:
            self._lib.pmx_wind_destroy(self._h)


class Wind:
    """The line-driven disk wind (2-D spherical) advanced on the GPU.

    ``advance_step()`` is the reference's AdvanceStep for this problem,
    ``split_source()`` its SplitSource (cooling), ``run(n)`` n iterations of
    the main loop with cooling, all on the device."""

    @classmethod
    def borrowed(cls, cfg: setups.WindConfig, handle) -> "Wind":
        """View of a context owned elsewhere (a batch's grid)."""
        w = cls.__new__(cls)
        w.cfg, w._h, w._lib, w._owned = cfg, handle, load_library(), False
        w.set_tstop(getattr(cfg, "tstop", 1.e30))
        return w

    def __init__(self, cfg: setups.WindConfig, fluxes=None, grid=None, sirocco=None):
        self.cfg = cfg
        self._owned = True
        lib = load_library()
        self._keep = []
        d = wind_desc(cfg, fluxes, grid, keep=self._keep, sirocco=sirocco)
        h = ctypes.c_void_p()
        _check(lib.pmx_wind_create(ctypes.byref(d), ctypes.byref(h)), "pmx_wind_create")
        self._keep = []
        self._h = h
        self._lib = lib
        self.set_tstop(getattr(cfg, "tstop", 1.e30))

    def close(self) -> None:
        """Destroy the context.  While a zero-copy snapshot view
        (snapshot_wait(copy=False)) is still referenced, the context (and
        the pinned buffer the view reads) is destroyed when the last such
        view is dropped."""
        if getattr(self, "_h", None):
            if self._owned:
                views = list(getattr(self, "_views", ()))
                if views:
                    _DeferredDestroy(self._lib, self._h, views)
                else:
                    self._lib.pmx_wind_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_prim(self, V: np.ndarray, init_heatcool: bool = True) -> None:
        """Upload V[6][1][nt][nr]; with init_heatcool the SIROCCO xi / T_r
        arrays are set from this state as the reference does at start-up."""
        V = np.ascontiguousarray(V, dtype=np.float64)
        if V.size != int(np.prod(self.cfg.shape)):
            raise PmxError(f"set_prim: {V.size} values, the grid holds {np.prod(self.cfg.shape)}")
        _check(self._lib.pmx_wind_set_prim(self._h, _dptr(V)), "set_prim")
        if init_heatcool:
            _check(self._lib.pmx_wind_set_heatcool(self._h, None, None), "set_heatcool")

    def set_heatcool(self, xi: np.ndarray, t_r: np.ndarray) -> None:
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        t_r = np.ascontiguousarray(t_r, dtype=np.float64)
        n = self.cfg.nx[0] * self.cfg.nx[1]
        if xi.size != n or t_r.size != n:
            raise PmxError(f"set_heatcool: xi and t_r need {n} values each")
        _check(self._lib.pmx_wind_set_heatcool(self._h, _dptr(xi), _dptr(t_r)), "set_heatcool")

    def set_prefactors(self, pre: np.ndarray | None) -> None:
        """Cooling prefactors [6][nt][nr] in prefactors.dat's column order
        (comp_h, comp_c, xray_h, brem_c, line_c, xi_ion), or None for 1.0."""
        if pre is None:
            _check(self._lib.pmx_wind_set_prefactors(self._h, None), "set_prefactors")
            return
        pre = np.asarray(pre, dtype=np.float64)
        n = self.cfg.nx[0] * self.cfg.nx[1]
        if pre.shape[0] not in (5, 6) or pre[0].size != n:
            raise PmxError("prefactors must be [5 or 6][nt][nr]")
        full = np.ones((6, n))                  # the library reads all six planes
        full[:pre.shape[0]] = pre.reshape(pre.shape[0], n)
        _check(self._lib.pmx_wind_set_prefactors(self._h, _dptr(full)), "set_prefactors")

    def set_sirocco(self, inp: "setups.SiroccoInputs") -> None:
        """The heat-cool arrays (after the reader's floors) and prefactors of
        a restarted SIROCCO cycle."""
        if inp.xi is not None:
            self.set_heatcool(*inp.heatcool_floored())
        self.set_prefactors(inp.pre)

    def get_prim(self) -> np.ndarray:
        out = np.empty(self.cfg.shape)
        _check(self._lib.pmx_wind_get_prim(self._h, _dptr(out)), "get_prim")
        return out

    def snapshot(self, user_vars: bool = True) -> None:
        """Queue an output snapshot (V and the production user variables);
        later steps may be queued before snapshot_wait()."""
        if any(True for _ in getattr(self, "_views", ())):
            raise PmxError("snapshot(): a zero-copy view of the previous snapshot "
                           "(snapshot_wait(copy=False)) is still referenced; the new "
                           "snapshot would overwrite it while it is read")
        self._snap_uv = bool(user_vars)
        _check(self._lib.pmx_wind_snapshot(self._h, int(user_vars)), "snapshot")

    def snapshot_wait(self, copy: bool = True) -> dict:
        """The last snapshot: V [6][1][nt][nr], uv [23][1][nt][nr] (or None),
        t, dt, step.  copy=False: read-only views of the pinned host copy;
        while any of them is referenced, snapshot() refuses (it would
        overwrite the buffer) and close() leaves the buffer to the views."""
        nt, nr = self.cfg.nx[1], self.cfg.nx[0]
        nv = 6 + (23 if self._snap_uv else 0)
        out = np.empty((nv, 1, nt, nr)) if copy else None
        t, dt, st = ctypes.c_double(), ctypes.c_double(), ctypes.c_long()
        _check(self._lib.pmx_wind_snapshot_wait(self._h, _dptr(out) if copy else None,
                                                ctypes.byref(t), ctypes.byref(dt),
                                                ctypes.byref(st)), "snapshot_wait")
        if not copy:
            ptr = ctypes.POINTER(ctypes.c_double)()
            _check(self._lib.pmx_wind_snapshot_host(self._h, ctypes.byref(ptr)), "snapshot_host")
            holder = _PinnedView(ptr, (nv, 1, nt, nr))
            if getattr(self, "_views", None) is None:
                self._views = weakref.WeakSet()
            self._views.add(holder)
            out = np.asarray(holder)
        return {"V": out[:6], "uv": out[6:] if self._snap_uv else None, "t": t.value,
                "dt": dt.value, "step": st.value}

    def get_cons(self) -> np.ndarray:
        out = np.empty(self.cfg.shape)
        _check(self._lib.pmx_wind_get_cons(self._h, _dptr(out)), "get_cons")
        return out

    def set_time(self, t: float, dt: float, step: int = 0) -> None:
        _check(self._lib.pmx_wind_set_time(self._h, t, dt, step), "set_time")

    def set_tstop(self, tstop: float) -> None:
        """runtime.tstop: run() cuts the step that reaches it and halts
        after it (PMX_FLAG_TSTOP)."""
        _check(self._lib.pmx_wind_set_tstop(self._h, tstop), "set_tstop")

    def get_time(self) -> dict:
        t, dt, inv, dtc = (ctypes.c_double() for _ in range(4))
        st, fl = ctypes.c_long(), ctypes.c_int()
        _check(self._lib.pmx_wind_get_time(self._h, ctypes.byref(t), ctypes.byref(dt),
                                           ctypes.byref(st), ctypes.byref(inv),
                                           ctypes.byref(dtc), ctypes.byref(fl)), "get_time")
        return {"t": t.value, "dt": dt.value, "step": st.value, "inv_dt_hyp": inv.value,
                "dt_cool": dtc.value, "flags": fl.value}

    def set_dts(self, inv_dt_hyp: float, dt_cool: float = 1.e38) -> None:
        _check(self._lib.pmx_wind_set_dts(self._h, inv_dt_hyp, dt_cool), "set_dts")

    def advance_step(self) -> None:
        _check(self._lib.pmx_wind_advance_step(self._h), "advance_step")

    def split_source(self) -> None:
        _check(self._lib.pmx_wind_split_source(self._h), "split_source")

    def run(self, nsteps: int, graph: bool = False) -> None:
        """nsteps iterations of the main loop; graph=True replays a captured
        HIP graph of 8 steps (bit-identical; a remainder runs eagerly)."""
        if graph:
            _check(self._lib.pmx_wind_run_graph(self._h, nsteps), "run_graph")
        else:
            _check(self._lib.pmx_wind_run(self._h, nsteps), "run")

    def sync(self) -> None:
        _check(self._lib.pmx_wind_sync(self._h), "sync")

    def set_surrogate(self, mlp: "Mlp | None", use_network: bool = True) -> None:
        """Surrogate-cooled SplitSource; use_network=False is the ablation
        that starts the same refinement without the network's guess."""
        self._mlp = mlp   # the context keeps a pointer to it
        _check(self._lib.pmx_wind_set_surrogate(self._h, mlp._h if mlp else None,
                                                0 if use_network else 1), "set_surrogate")

    def surrogate_stats(self, reset: bool = False) -> dict:
        a, b, c, d = (ctypes.c_longlong() for _ in range(4))
        _check(self._lib.pmx_wind_surrogate_stats(self._h, ctypes.byref(a), ctypes.byref(b),
                                                  ctypes.byref(c), ctypes.byref(d), int(reset)),
               "stats")
        return {"listed": a.value, "accepted": b.value, "fallback": c.value,
                "radiat_calls": d.value}

    def timing(self, enable: bool, reset: bool = True) -> dict:
        sm, cm = ctypes.c_double(), ctypes.c_double()
        sn, cn = ctypes.c_int(), ctypes.c_int()
        _check(self._lib.pmx_wind_timing(self._h, int(enable), int(reset), ctypes.byref(sm),
                                         ctypes.byref(sn), ctypes.byref(cm), ctypes.byref(cn)),
               "timing")
        return {"stage_ms": sm.value, "stage_launches": sn.value, "cool_ms": cm.value,
                "cool_launches": cn.value}


class WindBatch:
    """Independent wind grids advanced together (batched persistent stage
    kernel, batched cooling, optional HIP graph replay)."""

    def __init__(self, cfgs, sirocco=None):
        """``sirocco``: per grid, the setups.SiroccoInputs whose force
        multiplier fits a grid with krad = alpharad = 999 needs (or None)."""
        lib = load_library()
        self.cfgs = list(cfgs)
        sirocco = list(sirocco) if sirocco is not None else [None] * len(self.cfgs)
        keep = []
        descs = (WindDesc * len(self.cfgs))()
        for g, cfg in enumerate(self.cfgs):
            descs[g] = wind_desc(cfg, keep=keep, sirocco=sirocco[g])
        h = ctypes.c_void_p()
        _check(lib.pmx_wind_batch_create(descs, len(self.cfgs), ctypes.byref(h)), "batch_create")
        self._h, self._lib = h, lib
        self.grids = []
        for g, cfg in enumerate(self.cfgs):
            wh = ctypes.c_void_p()
            _check(lib.pmx_wind_batch_grid(h, g, ctypes.byref(wh)), "batch_grid")
            self.grids.append(Wind.borrowed(cfg, wh))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.pmx_wind_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def cells(self) -> int:
        c = ctypes.c_longlong()
        _check(self._lib.pmx_wind_batch_cells(self._h, ctypes.byref(c)), "cells")
        return c.value

    def set_step(self, step: int) -> None:
        _check(self._lib.pmx_wind_batch_set_step(self._h, step), "set_step")

    def set_surrogate(self, mlp: "Mlp | None", use_network: bool = True) -> None:
        self._mlp = mlp
        _check(self._lib.pmx_wind_batch_set_surrogate(self._h, mlp._h if mlp else None,
                                                      0 if use_network else 1), "set_surrogate")

    def autotune(self) -> dict:
        sh, bl, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(self._lib.pmx_wind_batch_autotune(self._h, ctypes.byref(sh), ctypes.byref(bl),
                                                 ctypes.byref(ms)), "autotune")
        return {"shape": sh.value, "blocks": bl.value, "ms_per_stage": ms.value}

    def set_shape(self, shape: int, blocks: int = 0) -> None:
        _check(self._lib.pmx_wind_batch_set_shape(self._h, shape, blocks), "set_shape")

    def run(self, nsteps: int, graph: bool = False) -> None:
        _check(self._lib.pmx_wind_batch_run(self._h, nsteps, int(graph)), "run")

    def sync(self) -> None:
        _check(self._lib.pmx_wind_batch_sync(self._h), "sync")

    def timing(self, enable: bool, reset: bool = True) -> dict:
        sm, cm = ctypes.c_double(), ctypes.c_double()
        sn, cn = ctypes.c_int(), ctypes.c_int()
        _check(self._lib.pmx_wind_batch_timing(self._h, int(enable), int(reset), ctypes.byref(sm),
                                               ctypes.byref(sn), ctypes.byref(cm),
                                               ctypes.byref(cn)), "timing")
        return {"stage_ms": sm.value, "stage_launches": sn.value, "cool_ms": cm.value,
                "cool_launches": cn.value}
