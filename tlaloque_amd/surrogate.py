"""Training and use of the learned cooling surrogate (host side).

The network (3 -> 64 -> 64 -> 1, GELU, bf16 weights, FP32 accumulation) maps
(log10 n, log10 T, log10(dt/t_cool)) to g = r / r1, where r = log10(T_new/T)
and r1 = -log10(e) d / (1 + d), d = dt/t_cool (``first_order``): the ratio
makes the network's error relative to the change of T.  Its training data
are generated on the GPU from the reference's cooling formula integrated to
convergence (cooling mode TLQ_COOL_SUBCYCLED), and it is trained on the GPU
by the library's own MFMA forward/backward kernels with Adam
(include/tlq.h: tlq_surrogate_*).  A trained context then runs the fused
step with cooling mode TLQ_COOL_SURROGATE, falling back per cell to the
exact integration (TLQ_COOL_INTEGRATED) where the network's inputs leave the training box or
its answer fails the per-cell check (bounds, and the exact cooling time from
T to T_new, which the piecewise power-law table gives in closed form, within
5% of the change of ln T of dt).
"""
from __future__ import annotations

import dataclasses
import os
import time
from typing import Optional

import numpy as np

from . import params as P
from .hydro import HydroCore

DEFAULT_SAMPLES = 1_000_000      # training set size of the cooling-medium configuration
DEFAULT_STEPS = 120000
DEFAULT_LR = 1e-2
DEFAULT_BATCH = 4096


@dataclasses.dataclass
class TrainReport:
    samples: int
    steps: int
    batch: int
    final_loss: float
    seconds: float
    heldout_rms: float      # dex of T_new, network vs converged integrator, held-out samples
    heldout_max: float      # dex
    heldout_rel_median: float  # median of |r_net - r| / |r|
    heldout_rel_p90: float     # 90th percentile of the same


def first_order(x2):
    """r1 = -log10(e) d / (1 + d), d = 10**x2 = dt/t_cool: the network's output scale."""
    d = 10.0 ** np.asarray(x2, dtype=np.float64)
    return -np.log10(np.e) * d / (1.0 + d)


def predict(core: HydroCore, x: np.ndarray) -> np.ndarray:
    """r = log10(T_new/T) of the trained network for inputs x (B, 3)."""
    return core.surrogate_eval(x).astype(np.float64) * first_order(x[:, 2])


def table_path() -> str:
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "Z1.0.dat")


def train(core: Optional[HydroCore] = None, samples: int = DEFAULT_SAMPLES,
          steps: int = DEFAULT_STEPS, batch: int = DEFAULT_BATCH, seed: int = 2024,
          heldout: int = 65536, lr: float = DEFAULT_LR) -> tuple:
    """Generate data and train on the GPU; returns (params, TrainReport).

    With no ``core`` a small scratch context (any mesh works: the network does
    not depend on it) is created with the Z1.0 cooling table."""
    own = core is None
    if own:
        p = P.HydroParams(nx=8, ny=8, nz=8, cooling=P.COOL_SUBCYCLED)
        core = HydroCore(p, cooling_table=P.load_cooling_table(table_path()))
    try:
        # held-out set: an independent draw, never seen in training
        core.surrogate_generate(heldout, seed=seed + 99)
        x, y = core.surrogate_samples(0, heldout)
        t0 = time.perf_counter()
        core.surrogate_generate(samples, seed=seed)
        core.surrogate_init(seed=seed + 1)
        loss = core.surrogate_train(steps, batch=batch, lr_max=lr, lr_min=0.01 * lr)
        theta = core.surrogate_params()
        el = time.perf_counter() - t0
        r1 = first_order(x[:, 2])
        g = core.surrogate_eval(x).astype(np.float64)
        err = np.abs(g - y) * np.abs(r1)
        rel = np.abs(g - y) / np.maximum(np.abs(y), 1e-30)
        rep = TrainReport(samples, steps, batch, loss, el, float(np.sqrt((err ** 2).mean())),
                          float(err.max()), float(np.median(rel)), float(np.percentile(rel, 90)))
        return theta, rep
    finally:
        if own:
            core.close()
