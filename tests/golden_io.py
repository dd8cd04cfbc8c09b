"""Loading of the committed golden fixtures (tests/golden/*.npz)."""
import json
import pathlib

import numpy as np

import setups

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def names(kind: str = "hydro"):
    """Fixture names: kind "hydro" (Cartesian step) or "wind".  States too
    large to store whole (*_sample: a subsample of cells) are not listed."""
    all_ = sorted(p.stem for p in GOLDEN.glob("*.npz") if not p.stem.endswith("_sample"))
    wind = [n for n in all_ if n.startswith("wind_")]
    return wind if kind == "wind" else [n for n in all_ if n not in wind]


def load(name):
    z = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    meta = json.loads(str(z["config"]))
    if "cooling" in meta:
        for k in ("nx", "r_range", "theta_range", "bc"):
            meta[k] = tuple(meta[k])
        cfg = setups.WindConfig(**meta)
        return cfg, {k: z[k] for k in z.files if k != "config"}
    meta["nx"] = tuple(meta["nx"])
    meta["bc"] = tuple(meta["bc"])
    meta["xbeg"] = tuple(meta["xbeg"])
    meta["xend"] = tuple(meta["xend"])
    cfg = setups.HydroConfig(**meta)
    return cfg, {k: z[k] for k in z.files if k != "config"}


def sirocco(g):
    """The SIROCCO restart inputs stored with a wind fixture (sir_* arrays)
    as a setups.SiroccoInputs, or None; and its start time (0 if none)."""
    keys = [k for k in g if k.startswith("sir_")]
    t0 = float(g["t0"]) if "t0" in g else 0.0
    if not keys:
        return None, t0
    return setups.SiroccoInputs(**{k[4:]: g[k] for k in keys}), t0
