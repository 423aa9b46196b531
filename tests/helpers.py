"""Shared test helpers: oracle plans for workloads, error metrics."""
import numpy as np

from listctrecons_amd import scan
from oracle import oracle as O


def oracle_prm(w):
    p = O.Prm()
    pp = w.prm()
    for f, _ in O.Prm._fields_:
        setattr(p, f, getattr(pp, f))
    return p


def oracle_plan(w):
    a, c = w.angles_couch()
    return O.plan_helical(oracle_prm(w), a, c), a, c


def rel_rmse(x, ref):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    m = np.isfinite(ref)
    return float(np.linalg.norm((x - ref)[m]) / np.linalg.norm(ref[m]))


def hu_errors(x, ref):
    """(RMSE, max-abs) in HU of the difference (reference-scale units)."""
    m = np.isfinite(ref)
    d = (np.asarray(x, np.float64) - np.asarray(ref, np.float64))[m] * scan.hu_per_unit()
    return float(np.sqrt(np.mean(d * d))), float(np.max(np.abs(d)))


def cpu_fan(w, noise=True):
    """Projection file of a workload from the FP64 oracle projector (+ numpy Poisson)."""
    pl, a, c = oracle_plan(w)
    proj = O.project_ellipsoids(pl.ctg, a, c, w.ellipsoids())
    if noise:
        rng = np.random.default_rng(w.noise_seed)
        n = rng.poisson(w.i0 * np.exp(-proj))
        proj = -np.log(np.maximum(n, 1) / w.i0)
    fan = np.zeros((w.data_offset + w.readings, w.nt, w.ns), np.float32)
    fan[w.data_offset:] = proj
    return fan
