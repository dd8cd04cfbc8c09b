"""Loading of the committed golden fixtures (tests/golden/*.npz)."""
import json
import pathlib

import numpy as np

import setups

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def names():
    return sorted(p.stem for p in GOLDEN.glob("*.npz"))


def load(name):
    z = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    meta = json.loads(str(z["config"]))
    meta["nx"] = tuple(meta["nx"])
    meta["bc"] = tuple(meta["bc"])
    meta["xbeg"] = tuple(meta["xbeg"])
    meta["xend"] = tuple(meta["xend"])
    cfg = setups.HydroConfig(**meta)
    return cfg, {k: z[k] for k in z.files if k != "config"}
