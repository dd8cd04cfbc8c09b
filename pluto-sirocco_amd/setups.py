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
HLLC, ROE = 0, 1
BC_PERIODIC, BC_OUTFLOW, BC_REFLECTIVE = range(3)

_BC_NAMES = {BC_PERIODIC: "periodic", BC_OUTFLOW: "outflow",
             BC_REFLECTIVE: "reflective"}
_SOLVER_NAMES = {HLLC: "hllc", ROE: "roe"}


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
                  "tstop            1.e30",
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
