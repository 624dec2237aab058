"""Shared test helpers: fixtures, parameter decoding, comparisons."""
from __future__ import annotations

import glob
import json
import os

import numpy as np

from tlaloque_amd import params as P

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE_PATH = os.path.join(ROOT, "tlaloque_amd", "data", "Z1.0.dat")


def cooling_table():
    return P.load_cooling_table(TABLE_PATH)


def params_from_json(s: str) -> P.HydroParams:
    d = json.loads(s)
    d["winds"] = [P.Wind(**w) for w in d.get("winds", [])]
    d["masses"] = [P.PointMass(**m) for m in d.get("masses", [])]
    d["bc"] = tuple(d["bc"])
    return P.HydroParams(**d)


def ref_cases():
    """The hydro reference runs (the self-gravity ones are ``sg_ref_cases``)."""
    return sorted(os.path.basename(f)[4:-4] for f in glob.glob(os.path.join(GOLDEN, "ref_*.npz"))
                  if not os.path.basename(f).startswith("ref_sg_"))


def sg_ref_cases():
    """The reference runs built with -DGRAV (self-gravity)."""
    return sorted(os.path.basename(f)[4:-4] for f in glob.glob(os.path.join(GOLDEN, "ref_sg_*.npz")))


def load_ref_sg(name: str):
    """(params, u0, prims, steps, phi0, phit) of a -DGRAV reference run."""
    z = np.load(os.path.join(GOLDEN, f"ref_{name}.npz"), allow_pickle=False)
    p = params_from_json(str(z["params"]))
    return p, z["u0"], z["prims"], [int(s) for s in z["steps"]], z["phi0"], z["phit"]


def load_ref(name: str):
    z = np.load(os.path.join(GOLDEN, f"ref_{name}.npz"), allow_pickle=False)
    p = params_from_json(str(z["params"]))
    return p, z["u0"], z["prims"], [int(s) for s in z["steps"]]


def interior(a: np.ndarray) -> np.ndarray:
    return a[2:-2, 2:-2, 2:-2]


def prims_from_cons(u: np.ndarray, gamma: float) -> np.ndarray:
    """calc_prims (source/HydroCore.f90:14-21) in numpy."""
    w = np.empty_like(u)
    w[..., 0] = u[..., 0]
    w[..., 1:4] = u[..., 1:4] / u[..., :1]
    pr = (gamma - 1.0) * (u[..., 4] - 0.5 * (u[..., 1] ** 2 + u[..., 2] ** 2 + u[..., 3] ** 2)
                          / u[..., 0])
    w[..., 4] = np.maximum(pr, 1e-20)
    return w


def max_rel(a: np.ndarray, b: np.ndarray) -> float:
    """Largest |a-b| relative to the field's scale (per component)."""
    worst = 0.0
    for e in range(a.shape[-1]):
        scale = np.abs(b[..., e]).max()
        if scale == 0:
            scale = 1.0
        worst = max(worst, float(np.abs(a[..., e] - b[..., e]).max() / scale))
    return worst


def l1_rel(a: np.ndarray, b: np.ndarray) -> float:
    """L1 = sum|a-b| / sum|b| (utils_analisis.jl:78-82)."""
    den = np.abs(b).sum()
    return float(np.abs(a - b).sum() / (den if den else 1.0))
