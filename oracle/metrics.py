"""Error metrics of the reference's analysis harness.

TEST INFRASTRUCTURE ONLY.  Restates ``_all_error_metrics`` and the norms of
utils_analisis.jl:23-94: the exact profile (xex, yex) is linearly
interpolated at each numerical abscissa inside its range, and

    L1   = sum |e| / sum |y_exact|             (utils_analisis.jl:62-69)
    L2   = sqrt(sum e^2 / sum y_exact^2)
    MAPE = mean relative error, RMSPE = rms relative error, Linf = max
           relative error (relative to |y_exact|, or to max|y_exact| where
           |y_exact| <= 1e-12).
"""
from __future__ import annotations

import numpy as np


def all_error_metrics(xex, yex, xnum, ynum):
    xex, yex = np.asarray(xex, float), np.asarray(yex, float)
    xnum, ynum = np.asarray(xnum, float), np.asarray(ynum, float)
    # range of numerical points covered by the exact table (:24-34)
    lo = int(np.argmax(xnum >= xex[0])) if xex[0] > xnum[0] else 0
    hi = int(np.nonzero(xnum <= xex[-1])[0][-1]) if xex[-1] < xnum[-1] else len(xnum) - 1
    scale = np.abs(yex).max()
    errs = []
    l1n = l1d = l2n = l2d = 0.0
    k = 1  # 0-based index of the right neighbour (the reference starts its search at 2)
    for i in range(lo, hi + 1):
        # _neighbor_search (:15-22): first exact abscissa > x, from the last one found
        while k < len(xex) and not (xex[k] > xnum[i]):
            k += 1
        if k >= len(xex):
            return None
        j = k - 1
        x1, y1, x2, y2 = xex[j], yex[j], xex[k], yex[k]
        m = (y2 - y1) / (x2 - x1)
        b = (x2 * y1 - y2 * x1) / (x2 - x1)
        yi = m * xnum[i] + b
        ea = abs(yi - ynum[i])
        er = ea / abs(yi) if abs(yi) > 1e-12 else ea / scale
        errs.append(er)
        l1n += ea
        l1d += abs(yi)
        l2n += ea * ea
        l2d += yi * yi
    errs = np.array(errs)
    return {"err": errs, "MAPE": float(errs.mean()), "RMSPE": float(np.sqrt((errs ** 2).mean())),
            "L1": l1n / l1d, "L2": float(np.sqrt(l2n / l2d)), "Linf": float(errs.max())}


def norm_l1(yex, ynum):
    """_norm_L1, utils_analisis.jl:78-82."""
    yex, ynum = np.asarray(yex, float), np.asarray(ynum, float)
    return float(np.abs(yex - ynum).sum() / np.abs(yex).sum())
