"""Workload definitions: solver configurations and seeded initial conditions.

A ``HydroConfig`` carries the run-time and compile-time choices that PLUTO
splits between ``definitions.h`` and ``pluto.ini`` (reconstruction, limiter,
Riemann solver, RK order, boundaries, CFL, grid).  Initial conditions are
generated here, once, as float64 arrays ``V[nv][k][j][i]`` (interior cells
only) and handed unchanged to both the GPU library and the CPU oracle, so
that both start from identical bits.

Noise comes from a counter-based SplitMix64 hash keyed by (seed, global
cell index), so any sub-block of a grid can be regenerated independently.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

NVAR = 5
RHO, VX1, VX2, VX3, PRS = range(5)

PLM, PPM = 0, 1
LIM_DEFAULT, LIM_VANLEER, LIM_MC, LIM_MINMOD, LIM_FLAT = range(5)
HLLC, ROE, HLL = 0, 1, 2   # HLL: the wind only (the production problem's [Solver])
BC_PERIODIC, BC_OUTFLOW, BC_REFLECTIVE = range(3)
# the production wind problem's userdef sides (cv_idl/init.c): X1-beg outflow
# with v_r <= 0, X2-beg reflective for velocities and outflow for rho and p
BC_WIND_INNER, BC_WIND_POLE = 3, 4

_BC_NAMES = {BC_PERIODIC: "periodic", BC_OUTFLOW: "outflow",
             BC_REFLECTIVE: "reflective", BC_WIND_INNER: "userdef", BC_WIND_POLE: "userdef"}
_SOLVER_NAMES = {HLLC: "hllc", ROE: "roe", HLL: "hll"}


@dataclasses.dataclass
class HydroConfig:
    dims: int
    nx: tuple                      # interior cells (nx1, nx2, nx3)
    recon: int = PLM
    limiter: int = LIM_VANLEER
    solver: int = HLLC
    rk: int = 2
    bc: tuple = (BC_PERIODIC,) * 6
    gamma: float = 5.0 / 3.0
    cfl: float = 0.4
    cfl_max_var: float = 1.1
    first_dt: float = 1.e-4
    xbeg: tuple = (-0.5, -0.5, -0.5)
    xend: tuple = (0.5, 0.5, 0.5)
    tstop: float = 1.e30          # pluto.ini [Time] tstop

    @property
    def nghost(self) -> int:
        return 3 if self.recon == PPM else 2

    @property
    def shape(self) -> tuple:
        """Interior array shape (nv, nz, ny, nx) of V."""
        nz = self.nx[2] if self.dims == 3 else 1
        return (NVAR, nz, self.nx[1], self.nx[0])

    @property
    def ncells(self) -> int:
        s = self.shape
        return s[1] * s[2] * s[3]

    def reference_problem(self) -> str:
        """Name of the oracle/_ref build whose definitions.h matches."""
        if self.recon == PPM:
            return f"cart{self.dims}d_ppm_rk{self.rk}"
        lim = {LIM_VANLEER: "vl", LIM_DEFAULT: "default", LIM_MC: "mc"}[self.limiter]
        return f"cart{self.dims}d_plm_{lim}_rk{self.rk}"

    def pluto_ini(self) -> str:
        """pluto.ini text equivalent to this configuration."""
        lines = ["[Grid]", ""]
        for d in range(3):
            n = self.nx[d] if d < self.dims else 1
            xb = self.xbeg[d] if d < self.dims else 0.0
            xe = self.xend[d] if d < self.dims else 1.0
            lines.append(f"X{d+1}-grid    1  {xb!r}  {n}  u  {xe!r}")
        lines += ["", "[Time]", "",
                  f"CFL              {self.cfl!r}",
                  f"CFL_max_var      {self.cfl_max_var!r}",
                  f"tstop            {float(self.tstop)!r}",
                  f"first_dt         {self.first_dt!r}",
                  "", "[Solver]", "",
                  f"Solver         {_SOLVER_NAMES[self.solver]}",
                  "", "[Boundary]", ""]
        labels = ["X1-beg", "X1-end", "X2-beg", "X2-end", "X3-beg", "X3-end"]
        for s, lab in enumerate(labels):
            lines.append(f"{lab}        {_BC_NAMES[self.bc[s]]}")
        lines += ["", "[Static Grid Output]", "",
                  "uservar    0",
                  "dbl       -1.0  -1   single_file",
                  "flt       -1.0  -1   single_file",
                  "vtk       -1.0  -1   single_file",
                  "tab       -1.0  -1",
                  "ppm       -1.0  -1",
                  "png       -1.0  -1",
                  "log        1000000",
                  "analysis  -1.0  1",
                  "", "[Parameters]", "",
                  f"GAMMA_EOS      {self.gamma!r}", ""]
        return "\n".join(lines)


# ----------------------------------------------------------------------
# counter-based noise
# ----------------------------------------------------------------------
def splitmix64(seed: int, idx: np.ndarray) -> np.ndarray:
    """Output number ``idx`` (0-based) of the SplitMix64 generator seeded
    with ``seed``: mix(seed + (idx + 1) * 0x9E3779B97F4A7C15 mod 2^64)."""
    golden = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (idx.astype(np.uint64) + np.uint64(1)) * golden
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, idx: np.ndarray) -> np.ndarray:
    """Uniform doubles in [0, 1) with 53 random bits."""
    return (splitmix64(seed, idx) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def cell_centers(cfg: HydroConfig, d: int) -> np.ndarray:
    """Cell centres of a uniform grid, as PLUTO builds them
    (Src/set_grid.c:396-403 and :213-215)."""
    n = cfg.nx[d] if d < cfg.dims else 1
    xl0, xr0 = (cfg.xbeg[d], cfg.xend[d]) if d < cfg.dims else (0.0, 1.0)
    dx = (xr0 - xl0) / float(n)
    out = np.empty(n)
    for i in range(n):
        xl = xl0 + float(i) * dx
        xr = xl + dx
        out[i] = 0.5 * (xl + xr)
    return out


def kelvin_helmholtz(cfg: HydroConfig, seed: int = 1234) -> np.ndarray:
    """2-D KH: rho=2, vx=+0.5 for |y|<0.25, else rho=1, vx=-0.5; p=2.5;
    vy = 0.01*(2u-1) with u from the counter hash."""
    nv, nz, ny, nx = cfg.shape
    y = cell_centers(cfg, 1)
    inner = (np.abs(y) < 0.25)[None, :, None]
    V = np.zeros(cfg.shape)
    V[RHO] = np.where(inner, 2.0, 1.0)
    V[VX1] = np.where(inner, 0.5, -0.5)
    idx = np.arange(nz * ny * nx, dtype=np.uint64).reshape(nz, ny, nx)
    V[VX2] = 0.01 * (2.0 * uniform01(seed, idx) - 1.0)
    V[VX3] = 0.0
    V[PRS] = 2.5
    return V


def sedov(cfg: HydroConfig, seed: int = 1234, energy: float = 1.0,
          radius_cells: float = 3.0, p0: float = 1.e-5,
          noise: float = 0.01) -> np.ndarray:
    """3-D Sedov blast: rho = exp(noise * N(0,1)) (lognormal), p = p0, and
    energy ``energy`` deposited as pressure (gamma-1) E / Vol in the cells
    whose centre lies within ``radius_cells`` cell widths of the domain
    centre; Vol is the summed volume of those cells."""
    nv, nz, ny, nx = cfg.shape
    xs = [cell_centers(cfg, d) for d in range(3)]
    dxs = [(cfg.xend[d] - cfg.xbeg[d]) / cfg.nx[d] for d in range(cfg.dims)]
    ctr = [0.5 * (cfg.xbeg[d] + cfg.xend[d]) for d in range(cfg.dims)]
    r2 = (xs[0][None, None, :] - ctr[0]) ** 2 + (xs[1][None, :, None] - ctr[1]) ** 2
    if cfg.dims == 3:
        r2 = r2 + (xs[2][:, None, None] - ctr[2]) ** 2
    rad = radius_cells * dxs[0]
    inside = r2 <= rad * rad
    cellvol = float(np.prod(dxs))
    vol = float(inside.sum()) * cellvol
    idx = np.arange(nz * ny * nx, dtype=np.uint64).reshape(nz, ny, nx)
    u1 = uniform01(seed, 2 * idx)
    u2 = uniform01(seed, 2 * idx + np.uint64(1))
    z = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)
    V = np.zeros(cfg.shape)
    V[RHO] = np.exp(noise * z)
    V[PRS] = np.where(inside, (cfg.gamma - 1.0) * energy / vol, p0)
    return V


def c1_kh(n: int = 256) -> HydroConfig:
    """2-D KH, PLM(van Leer) + HLLC + RK2, CFL 0.4, periodic, 256^2."""
    return HydroConfig(dims=2, nx=(n, n, 1), recon=PLM, limiter=LIM_VANLEER,
                       solver=HLLC, rk=2, bc=(BC_PERIODIC,) * 6, cfl=0.4,
                       first_dt=1.e-4)


def c2_sedov(n: int = 256) -> HydroConfig:
    """3-D Sedov, PPM + HLLC + RK3, CFL 0.3, outflow, 256^3."""
    return HydroConfig(dims=3, nx=(n, n, n), recon=PPM, limiter=LIM_DEFAULT,
                       solver=HLLC, rk=3, bc=(BC_OUTFLOW,) * 6, cfl=0.3,
                       first_dt=1.e-6)


# ----------------------------------------------------------------------
# Line-driven disk wind (2-D spherical r-theta)
# ----------------------------------------------------------------------
NVAR_WIND = 6                 # rho, v_r, v_theta, v_phi, p, tracer
TRC = 5
CONST = {  # Src/pluto.h:296-321
    "amu": 1.66053886e-24, "c": 2.99792458e10, "G": 6.6726e-8, "kB": 1.3806505e-16,
    "mp": 1.67262171e-24, "Msun": 2.e33, "PI": 3.14159265358979, "sigmaT": 6.6524e-25,
}
COOL_NONE, COOL_BLONDIN, COOL_PLAIN = 0, 1, 2
_COOL_NAMES = {COOL_BLONDIN: "BLONDIN", COOL_PLAIN: "TABULATED"}


@dataclasses.dataclass
class WindConfig:
    """2-D spherical line-driven wind: PLM (van Leer, general-grid
    coefficients) + HLLC + RK2, point-mass gravity, SIROCCO-mode line force
    with 36 angle bins, one passive tracer, and a split cooling step
    (BLONDIN root solve or the sub-cycled Radiat() ODE integrator)."""
    nx: tuple = (512, 256)                   # (N_r, N_theta)
    r_range: tuple = (1.0, 100.0)            # code units, logarithmic grid
    theta_range: tuple = (0.0, 0.5 * 3.14159265358979)
    cooling: int = COOL_PLAIN
    solver: int = HLLC
    gamma: float = 5.0 / 3.0
    cfl: float = 0.4
    cfl_max_var: float = 1.1
    first_dt: float = 1.e-4
    bc: tuple = (BC_OUTFLOW, BC_OUTFLOW, BC_REFLECTIVE, BC_REFLECTIVE)
    mu: float = 0.6
    t_iso: float = 1.e5
    l_star: float = 1.e38
    f_x: float = 0.1
    f_uv: float = 0.1
    t_x: float = 1.6e5
    krad: float = 0.2
    alpharad: float = -0.6
    cent_mass: float = 0.6 * 2.e33
    nangles: int = 36
    # > 0: the production problem's internal boundary (UserDefBoundary side 0:
    # density / pressure / temperature floors), density floor in g/cm^3
    dfloor: float = 0.0
    # SHOCK_FLATTENING MULTID: shocked zones take minmod slopes and HLL fluxes
    shock_flattening: bool = False
    # CHAR_LIMITING YES: PLM limits the characteristic differences
    char_limiting: bool = False
    # ENTROPY_SWITCH ALWAYS: the pressure comes from an advected entropy
    entropy_switch: bool = False
    # units of the problem (definitions.h)
    unit_density: float = 1e-14
    unit_length: float = 1e9
    unit_velocity: float = 1e9
    tstop: float = 1.e30                     # pluto.ini [Time] tstop
    # line force's power per angle bin: 0 the reference's form, 1 factored
    # (one exp2/log2 per bin for both LineForce calls; a few 1e-15 off)
    force_power: int = 0
    # the cv_idl disk (its midplane row): RHO_0 [g/cm^3], R_0, RHO_ALPHA, DISK_MDOT [g/s]
    rho_0: float = 1.e-9
    r_0: float = 1.0
    rho_alpha: float = 2.0
    disk_mdot: float = 6.3e17

    dims = 2
    nghost = 2

    @property
    def shape(self) -> tuple:
        return (NVAR_WIND, 1, self.nx[1], self.nx[0])

    @property
    def ncells(self) -> int:
        return self.nx[0] * self.nx[1]

    def reference_problem(self) -> str:
        if self.dfloor > 0.0:
            if self.cooling != COOL_PLAIN:
                raise ValueError("the cv_idl boundary build uses the plain cooling")
            if self.entropy_switch:
                if not (self.shock_flattening and self.char_limiting):
                    raise ValueError("the entropy switch is built with the production "
                                     "reconstruction")
                return "wind_cvidl_prod"
            if self.char_limiting:
                if not self.shock_flattening:
                    raise ValueError("characteristic limiting is built with shock flattening")
                return "wind_cvidl_char"
            return "wind_cvidl_flat" if self.shock_flattening else "wind_cvidl"
        if self.shock_flattening or self.char_limiting or self.entropy_switch:
            raise ValueError("shock flattening / characteristic limiting are built with the "
                             "cv_idl boundaries only")
        return {COOL_NONE: "wind_nocool", COOL_BLONDIN: "wind_blondin",
                COOL_PLAIN: "wind_plain"}[self.cooling]

    def pluto_ini(self, output: bool = False) -> str:
        """pluto.ini of the reference build; ``output``: the production
        problem's user variables and a .dbl file every step."""
        lines = ["[Grid]", "",
                 f"X1-grid    1  {self.r_range[0]!r}  {self.nx[0]}  l+  {self.r_range[1]!r}",
                 f"X2-grid    1  {self.theta_range[0]!r}  {self.nx[1]}  u  "
                 f"{self.theta_range[1]!r}",
                 "X3-grid    1  0.0  1  u  1.0",
                 "", "[Time]", "",
                 f"CFL              {self.cfl!r}",
                 f"CFL_max_var      {self.cfl_max_var!r}",
                 f"tstop            {float(self.tstop)!r}",
                 f"first_dt         {self.first_dt!r}",
                 "", "[Solver]", "", f"Solver         {_SOLVER_NAMES[self.solver]}",
                 "", "[Boundary]", ""]
        labels = ["X1-beg", "X1-end", "X2-beg", "X2-end"]
        for s, lab in enumerate(labels):
            lines.append(f"{lab}        {_BC_NAMES[self.bc[s]]}")
        lines += ["X3-beg        outflow", "X3-end        outflow",
                  "", "[Static Grid Output]", "",
                  f"uservar    {len(USER_VARS)} {' '.join(USER_VARS)}" if output else "uservar    0",
                  "dbl       -1.0   1   single_file" if output else "dbl       -1.0  -1   single_file",
                  "flt       -1.0  -1   single_file",
                  "vtk       -1.0  -1   single_file",
                  "tab       -1.0  -1", "ppm       -1.0  -1", "png       -1.0  -1",
                  "log        1000000", "analysis  -1.0  1",
                  "", "[Parameters]", ""]
        if self.dfloor > 0.0:
            # the cv_idl parameter list, in its definitions.h order (RHO_0,
            # RHO_ALPHA and DISK_MDOT enter its midplane row)
            lines += [f"MU             {self.mu!r}", f"RHO_0          {self.rho_0!r}",
                      f"R_0            {self.r_0!r}", f"RHO_ALPHA      {self.rho_alpha!r}",
                      f"CENT_MASS      {self.cent_mass!r}",
                      f"DISK_MDOT      {self.disk_mdot!r}", f"T_ISO          {self.t_iso!r}",
                      f"L_star         {self.l_star!r}", f"f_x            {self.f_x!r}",
                      f"f_uv           {self.f_uv!r}", f"T_x            {self.t_x!r}",
                      f"KRAD           {self.krad!r}", f"ALPHARAD       {self.alpharad!r}",
                      f"DFLOOR         {self.dfloor!r}", f"GAMMA          {self.gamma!r}", ""]
        else:
            lines += [f"MU             {self.mu!r}", f"T_ISO          {self.t_iso!r}",
                      f"L_star         {self.l_star!r}", f"f_x            {self.f_x!r}",
                      f"f_uv           {self.f_uv!r}", f"T_x            {self.t_x!r}",
                      f"KRAD           {self.krad!r}", f"ALPHARAD       {self.alpharad!r}",
                      f"CENT_MASS      {self.cent_mass!r}", f"GAMMA_EOS      {self.gamma!r}", ""]
        return "\n".join(lines)


def make_grid(kind: str, xL: float, xR: float, n: int, ng: int):
    """Cell faces, centres and widths of one direction including ng ghost
    cells, built like the reference (Src/set_grid.c: MakeGrid 'u' and 'l+'
    segments, ghost cells copying the adjacent width, x = (xl + xr)/2).
    Scalar libm arithmetic (math module) keeps the bits identical."""
    nt = n + 2 * ng
    xl = [0.0] * (nt + 1)
    xr = [0.0] * nt
    dx = [0.0] * nt
    ib, ie = ng, ng + n - 1
    if kind == "u":
        for i in range(ib, ie + 1):
            dx[i] = (xR - xL) / float(n)
            xl[i] = xL + float(i - ib) * dx[i]
            xr[i] = xl[i] + dx[i]
    elif kind == "l+":
        dy = math.log10((xR + abs(xL) - xL) / abs(xL))
        dy /= float(n)
        xl[ib] = xL
        for i in range(ib, ie + 1):
            dx[i] = (xl[i] + abs(xL) - xL) * (math.pow(10.0, dy) - 1.0)
            xr[i] = xl[i] + dx[i]
            xl[i + 1] = xr[i]
    else:
        raise ValueError(kind)
    for i in range(ng):
        dx[i] = dx[ib]
        xl[i] = xl[ib] - (ng - i) * dx[ib]
        xr[i] = xl[i] + dx[ib]
        dx[ie + i + 1] = dx[ie]
        xl[ie + i + 1] = xl[ie] + (i + 1) * dx[ie]
        xr[ie + i + 1] = xl[ie] + (i + 2) * dx[ie]
    xl = xl[:nt]
    x = [0.5 * (xl[i] + xr[i]) for i in range(nt)]
    return {"x": np.array(x), "xl": np.array(xl), "xr": np.array(xr), "dx": np.array(dx)}


def wind_grid(cfg: WindConfig):
    return (make_grid("l+", cfg.r_range[0], cfg.r_range[1], cfg.nx[0], cfg.nghost),
            make_grid("u", cfg.theta_range[0], cfg.theta_range[1], cfg.nx[1], cfg.nghost))


def wind_ic(cfg: WindConfig, seed: int = 1234, rho0_cgs: float = 1.e-12,
            temp: float = 1.e5, noise: float = 0.01) -> np.ndarray:
    """rho = rho0 r^-2 (1 + noise (2u-1)), T = temp (p = rho T / (KELVIN mu)),
    Keplerian rotation v_phi = sqrt(G M / r) sin(theta), tracer = 1."""
    gr, gt = wind_grid(cfg)
    g = cfg.nghost
    r = gr["x"][g:-g][None, None, :]
    th = gt["x"][g:-g][None, :, None]
    kelvin = cfg.unit_velocity * cfg.unit_velocity * CONST["amu"] / CONST["kB"]
    nz, ny, nx = 1, cfg.nx[1], cfg.nx[0]
    idx = np.arange(ny * nx, dtype=np.uint64).reshape(1, ny, nx)
    V = np.zeros(cfg.shape)
    V[RHO] = rho0_cgs / cfg.unit_density * r ** -2 * (1.0 + noise * (2.0 * uniform01(seed, idx) - 1.0))
    gm_code = CONST["G"] * cfg.cent_mass / (cfg.unit_length * cfg.unit_velocity ** 2)
    V[VX3] = np.sqrt(gm_code / r) * np.sin(th) * np.ones_like(V[RHO])
    V[PRS] = V[RHO] * temp / (kelvin * cfg.mu)
    V[TRC] = 1.0
    return V


def wind_ic_mixed(cfg: WindConfig, seed: int = 77, rho0_cgs: float = 1.e-10,
                  rho_decades: float = 2.0, t_range: tuple = (3.e3, 3.e7),
                  v_cgs: float = 3.e7) -> np.ndarray:
    """Stress state for the wind step: density rho0 r^-2 times a log-uniform
    factor spanning +-rho_decades/2 decades, temperature log-uniform over
    t_range (cells below 1e4 K and stiffly cooling cells both occur),
    random poloidal velocities up to v_cgs, Keplerian rotation, tracer in
    {0, 1} so the Radiat() tracer gate takes both branches."""
    gr, gt = wind_grid(cfg)
    g = cfg.nghost
    r = gr["x"][g:-g][None, None, :]
    th = gt["x"][g:-g][None, :, None]
    kelvin = cfg.unit_velocity * cfg.unit_velocity * CONST["amu"] / CONST["kB"]
    ny, nx = cfg.nx[1], cfg.nx[0]
    n = ny * nx
    idx = np.arange(n, dtype=np.uint64).reshape(1, ny, nx)
    u = [uniform01(seed + k, idx) for k in range(5)]
    V = np.zeros(cfg.shape)
    V[RHO] = rho0_cgs / cfg.unit_density * r ** -2 * 10.0 ** (rho_decades * (u[0] - 0.5))
    lt0, lt1 = np.log10(t_range[0]), np.log10(t_range[1])
    T = 10.0 ** (lt0 + (lt1 - lt0) * u[1])
    V[PRS] = V[RHO] * T / (kelvin * cfg.mu)
    V[VX1] = v_cgs / cfg.unit_velocity * (2.0 * u[2] - 1.0)
    V[VX2] = v_cgs / cfg.unit_velocity * (2.0 * u[3] - 1.0)
    gm_code = CONST["G"] * cfg.cent_mass / (cfg.unit_length * cfg.unit_velocity ** 2)
    V[VX3] = np.sqrt(gm_code / r) * np.sin(th) * np.ones_like(V[RHO])
    V[TRC] = (u[4] > 0.1).astype(np.float64)
    return V


def wind_fluxes(cfg: WindConfig):
    """Synthetic 36-bin UV flux field (cgs) standing in for SIROCCO's
    directional_flux_{r,theta,phi}.dat: a central source L_uv/(4 pi r^2)
    spread over all bins with weight ~ max(cos(theta_a - theta), 0)^2 + 0.05,
    theta_a = (a + 1/2) 2 pi / 36.  Returns (flux_r, flux_t, flux_p), each
    [nangles][1][N_theta][N_r]."""
    gr, gt = wind_grid(cfg)
    g = cfg.nghost
    r = gr["x"][g:-g][None, :] * cfg.unit_length
    th = gt["x"][g:-g][:, None]
    luv = cfg.l_star * cfg.f_uv
    f0 = luv / (4.0 * CONST["PI"] * r * r)
    a = (np.arange(cfg.nangles) + 0.5) * (2.0 * CONST["PI"]) / cfg.nangles
    w = np.maximum(np.cos(a[:, None, None] - th[None]), 0.0) ** 2 + 0.05
    w = w / w.sum(axis=0, keepdims=True)
    fr = f0[None] * w * np.cos(a[:, None, None] - th[None])
    ft = f0[None] * w * np.sin(a[:, None, None] - th[None])
    fp = np.zeros_like(fr)
    return fr[:, None], ft[:, None], fp[:, None]


def write_flux_files(cfg: WindConfig, directory, fluxes=None) -> None:
    """Write directional_flux_{r,theta,phi}.dat in the format the reference
    parses (Src/LineDriven/line_connect.c:84-165)."""
    import pathlib
    directory = pathlib.Path(directory)
    fr, ft, fp = fluxes if fluxes is not None else wind_fluxes(cfg)
    gr, gt = wind_grid(cfg)
    g = cfg.nghost
    xs, ts = gr["x"][g:-g], gt["x"][g:-g]
    for name, arr in (("r", fr), ("theta", ft), ("phi", fp)):
        with open(directory / f"directional_flux_{name}.dat", "w") as f:
            f.write(f"# synthetic directional flux ({name})\n# nangles {cfg.nangles}\n")
            for j in range(cfg.nx[1]):
                for i in range(cfg.nx[0]):
                    vals = " ".join(f"{v:.17e}" for v in arr[:, 0, j, i])
                    f.write(f"{i} {j} 0 {xs[i] * cfg.unit_length:.17e} {ts[j]:.17e} {vals}\n")


# The production problem's user variables, in its uservar order
# (cv_idl/pluto_sirocco_sub.py:86), as ComputeUserVar fills them
# (cv_idl/userdef_output.c:20-260).
USER_VARS = ("XI", "T", "T_r", "ch", "cc", "lc", "bc", "xh", "ch_pre", "cc_pre", "lc_pre",
             "bc_pre", "xh_pre", "XI_pre", "ne", "nh", "gr", "gt", "gp", "dv_ds", "t",
             "M_max1", "M_max2")


@dataclasses.dataclass
class SiroccoInputs:
    """What a SIROCCO cycle hands PLUTO besides the fluxes, as its files
    hold them (Src/LineDriven/line_connect.c:183-497):

    * ``xi``, ``t_r`` [nt][nr]: py_heatcool.dat's ionisation parameter and
      radiation temperature (the reader floors them at 1 and 1e3 K);
    * ``pre`` [6][nt][nr]: prefactors.dat's comp_h, comp_c, xray_h, brem_c,
      line_c and xi_ion prefactors;
    * ``fm_t`` [points], ``fm_m`` [points][nt][nr]: M_UV_data.dat's force
      multiplier fits M(t) per cell (used when krad = alpharad = 999).

    A field left None: that file is not written."""
    xi: np.ndarray | None = None
    t_r: np.ndarray | None = None
    pre: np.ndarray | None = None
    fm_t: np.ndarray | None = None
    fm_m: np.ndarray | None = None

    def heatcool_floored(self):
        """xi and t_r as read_sirocco_heatcool leaves them (line_connect.c:354-355)."""
        xi, t_r = self.xi.copy(), self.t_r.copy()
        xi[xi < 1.0] = 1.0
        t_r[t_r < 1.e3] = 1.e3
        return xi, t_r


def sirocco_inputs(cfg: WindConfig, V: np.ndarray, seed: int = 5, heatcool: bool = True,
                   prefactors: bool = True, fm_points: int = 0,
                   fm_range: tuple = (1.e-7, 1.e3)) -> SiroccoInputs:
    """Synthetic SIROCCO cycle outputs for the wind state V: xi spread a
    decade either side of the start-up value L_x / (n_H r^2) and T_r within
    a factor 2 of T_x (3% of cells below each reader floor); prefactors in
    [0.5, 1.5); fits M(t) = k t^alpha (1 + 0.3 (u - 1/2)) with per-cell
    k in [0.1, 0.4) and alpha in [-0.7, -0.5) at ``fm_points`` values of t
    spaced logarithmically over ``fm_range``."""
    nt, nr = cfg.nx[1], cfg.nx[0]
    idx = np.arange(nt * nr, dtype=np.uint64)
    u = [uniform01(seed + s, idx).reshape(nt, nr) for s in range(12)]
    out = SiroccoInputs()
    if heatcool:
        gr, _ = wind_grid(cfg)
        g = cfg.nghost
        r = gr["x"][g:-g][None, :] * cfg.unit_length
        nH = V[RHO, 0] * cfg.unit_density / (1.43 * CONST["mp"])
        xi0 = cfg.l_star * cfg.f_x / nH / (r * r)
        out.xi = xi0 * 10.0 ** (2.0 * u[0] - 1.0)
        out.xi[u[1] < 0.03] = 0.5
        out.t_r = cfg.t_x * 2.0 ** (2.0 * u[2] - 1.0)
        out.t_r[u[3] < 0.03] = 500.0
    if prefactors:
        out.pre = np.stack([0.5 + u[4 + q] for q in range(6)])
    if fm_points:
        out.fm_t = np.logspace(np.log10(fm_range[0]), np.log10(fm_range[1]), fm_points)
        k = 0.1 + 0.3 * u[10]
        alpha = -0.7 + 0.2 * u[11]
        wig = np.stack([uniform01(seed + 100 + q, idx).reshape(nt, nr) for q in range(fm_points)])
        out.fm_m = k[None] * out.fm_t[:, None, None] ** alpha[None] * (1.0 + 0.3 * (wig - 0.5))
    return out


def write_sirocco_files(cfg: WindConfig, directory, inp: SiroccoInputs,
                        V: np.ndarray | None = None) -> None:
    """Write py_heatcool.dat, prefactors.dat and M_UV_data.dat for the
    fields of ``inp`` that are set, in the layouts the reference parses
    (line_connect.c:183-257, :330-343, :430-437); every cell of the grid
    gets one record, values with 17 significant digits.  V (for the
    density columns) defaults to a unit density."""
    import pathlib
    directory = pathlib.Path(directory)
    gr, gt = wind_grid(cfg)
    g = cfg.nghost
    xs, ts = gr["x"][g:-g], gt["x"][g:-g]
    nt, nr = cfg.nx[1], cfg.nx[0]
    dens = (V[RHO, 0] if V is not None else np.ones((nt, nr))) * cfg.unit_density
    e = lambda v: f"{v:.17e}"
    if inp.xi is not None:
        with open(directory / "py_heatcool.dat", "w") as f:
            f.write("i j rcen thetacen vol t_e t_r xi ne heat_xray heat_comp heat_lines "
                    "heat_ff cool_comp cool_lines cool_ff rho n_h\n")
            for j in range(nt):
                for i in range(nr):
                    rest = " ".join(["1.0e+30", "1.0e+05", e(inp.t_r[j, i]), e(inp.xi[j, i]),
                                     "1.0e+00"] + ["0.0e+00"] * 7 +
                                    [e(dens[j, i]), e(dens[j, i] / (1.43 * CONST["mp"]))])
                    f.write(f"{i} {j} {e(xs[i] * cfg.unit_length)} {e(ts[j])} {rest}\n")
    if inp.pre is not None:
        with open(directory / "prefactors.dat", "w") as f:
            f.write("i rcen j thetacen rho comp_h comp_c xray_h brem_c line_c xi_ion\n")
            for j in range(nt):
                for i in range(nr):
                    vals = " ".join(e(inp.pre[q, j, i]) for q in range(6))
                    f.write(f"{i} {e(xs[i] * cfg.unit_length)} {j} {e(ts[j])} "
                            f"{e(dens[j, i])} {vals}\n")
    if inp.fm_t is not None:
        n = len(inp.fm_t)
        with open(directory / "M_UV_data.dat", "w") as f:
            f.write(f"MPOINTS {n}\n")
            f.write("t " + " ".join(e(t) for t in inp.fm_t) + "\n")
            for j in range(nt):
                for i in range(nr):
                    vals = " ".join(e(inp.fm_m[q, j, i]) for q in range(n))
                    f.write(f"{i} {j} {e(xs[i] * cfg.unit_length)} {e(ts[j])} {vals}\n")


def c3_wind(nr: int = 512, nt: int = 256, cooling: int = COOL_PLAIN) -> WindConfig:
    return WindConfig(nx=(nr, nt), cooling=cooling)


def cvidl_wind(nr: int = 512, nt: int = 256, dfloor: float = 1.e-15,
               cooling: int = COOL_PLAIN) -> WindConfig:
    """The wind with the production problem's boundaries
    (Test_Problems/LineDrivenWind/cv_idl: X1-beg and X2-beg userdef,
    X1-end outflow, X2-end reflective) and its internal-boundary floors."""
    return WindConfig(nx=(nr, nt), cooling=cooling, dfloor=dfloor,
                      bc=(BC_WIND_INNER, BC_OUTFLOW, BC_WIND_POLE, BC_REFLECTIVE))


def c5_batch(n_grids: int = 64, seed: int = 2024, cooling: int = COOL_PLAIN,
             size_range=((128, 64), (768, 384))) -> list:
    """Parameter sweep of independent wind grids: N_r uniform over
    [128, 768] and N_theta uniform over [64, 384] (multiples of 8), L_x log-
    uniform in 1e36..1e38 erg/s (L_star = L_x / f_x) and the force-multiplier
    k uniform in 0.1..0.4, all drawn from the seeded counter-based generator."""
    (r0, t0), (r1, t1) = size_range
    idx = np.arange(4 * n_grids, dtype=np.uint64)
    u = uniform01(seed, idx).reshape(n_grids, 4)
    cfgs = []
    for g in range(n_grids):
        nr = int(r0 + (r1 - r0) * u[g, 0]) // 8 * 8
        nt = int(t0 + (t1 - t0) * u[g, 1]) // 8 * 8
        lx = 10.0 ** (36.0 + 2.0 * u[g, 2])
        cfg = WindConfig(nx=(max(nr, r0), max(nt, t0)), cooling=cooling)
        cfgs.append(dataclasses.replace(cfg, l_star=lx / cfg.f_x, krad=0.1 + 0.3 * u[g, 3]))
    return cfgs
