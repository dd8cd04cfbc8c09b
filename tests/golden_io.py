"""Loading of the committed golden fixtures (tests/golden/*.npz)."""
import json
import pathlib

import numpy as np

import setups

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def names(kind: str = "hydro"):
    """Fixture names: kind "hydro" (Cartesian step) or "wind"."""
    all_ = sorted(p.stem for p in GOLDEN.glob("*.npz"))
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
