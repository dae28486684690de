"""Calibration of the surrogate's fallback threshold.

    python -m pmmoto_amd.surrogate.calibrate [--n 512] [--seed0 7000000] [--out calibration_128.json]

On n calibration packs (seeds seed0 + i; the training packs are 5_000_000 +
i and the held-out ones 6_000_000 + i) the network's sigma and each curve's
largest absolute error against the exact sweep are computed.  For each
target error the threshold tau is the largest sigma at which every
calibration curve with sigma <= tau is within the target at all of its
points (split-conformal in spirit: a curve is returned only where the
calibration set shows no curve of that sigma or less beyond the target).
A target that the curve of the smallest sigma already misses gets a tau
below every sigma: every domain falls back.

The file also records how well sigma tracks the error (Spearman rank
correlation, the error quantiles) and the porosity range of the training
distribution; SurrogatePorosimetry sends domains outside that range to the
exact sweep whatever their sigma.
"""

from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

CALIBRATION = os.path.join(os.path.dirname(__file__), "calibration_128.json")
TARGETS = (0.02, 0.03, 0.05, 0.08)
POROSITY_RANGE = (0.25, 0.6)  # the training draw (pmmoto_amd.surrogate.data)


def curve_errors(pred: np.ndarray, exact: np.ndarray) -> np.ndarray:
    """Largest absolute error of each curve."""
    return np.max(np.abs(np.asarray(pred, np.float64) - np.asarray(exact, np.float64)), axis=1)


def choose_tau(sigma: np.ndarray, err: np.ndarray, target: float):
    """The threshold that keeps every curve with sigma below the smallest
    sigma of a curve beyond `target` (None: no curve is beyond it, keep
    all; 0.0: the curve of the smallest sigma is, keep none)."""
    sigma = np.asarray(sigma, np.float64)
    bad = sigma[np.asarray(err) > target]
    if bad.size == 0:
        return None
    return float(max(0.0, np.nextafter(bad.min(), -np.inf)))


def spearman(a: np.ndarray, b: np.ndarray) -> float:
    ra = np.argsort(np.argsort(a)).astype(np.float64)
    rb = np.argsort(np.argsort(b)).astype(np.float64)
    return float(np.corrcoef(ra, rb)[0, 1])


def report(sigma: np.ndarray, err: np.ndarray, taus: dict) -> dict:
    """How sigma and the per-curve error relate on one set, and what the
    thresholds keep."""
    out = {
        "curves": int(len(err)),
        "max_error_quantiles": {q: round(float(np.quantile(err, float(q))), 5) for q in ("0.5", "0.9", "0.99", "1")},
        "sigma_quantiles": {q: round(float(np.quantile(sigma, float(q))), 5) for q in ("0", "0.5", "0.9", "1")},
        "spearman_sigma_vs_max_error": round(spearman(sigma, err), 3),
        "fraction_max_error_above": {str(t): round(float(np.mean(err > t)), 4) for t in TARGETS},
        "at_tau": {},
    }
    for t, tau in taus.items():
        keep = np.ones(len(sigma), bool) if tau is None else sigma <= tau
        out["at_tau"][t] = {
            "tau": tau,
            "fallback_fraction": round(float(1.0 - keep.mean()), 4),
            "max_error_returned": round(float(err[keep].max()), 5) if keep.any() else None,
        }
    return out


def load_calibration(path: str = CALIBRATION) -> dict | None:
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def tau_for(target: float, path: str = CALIBRATION) -> float:
    """The calibrated threshold of `target` (one of the calibrated targets;
    inf when no calibration curve missed it)."""
    cal = load_calibration(path)
    if cal is None:
        raise FileNotFoundError(f"no surrogate calibration at {path}")
    taus = cal["tau"]
    if str(target) not in taus:
        raise ValueError(f"target {target} not calibrated; calibrated: {sorted(taus)}")
    tau = taus[str(target)]
    return float("inf") if tau is None else float(tau)


def main(argv=None) -> int:
    from .data import make_dataset
    from .train import load

    ap = argparse.ArgumentParser()
    ap.add_argument("--weights", default=os.path.join(os.path.dirname(__file__), "weights_128.npz"))
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--seed0", type=int, default=7_000_000)
    ap.add_argument("--out", default=CALIBRATION)
    a = ap.parse_args(argv)
    net = load(a.weights)
    x, exact, _ = make_dataset(a.n, a.seed0, size=net.size)
    curve, sigma = net.predict(x, batch=min(512, a.n))
    sigma = sigma.cpu().numpy().astype(np.float64)
    err = curve_errors(curve.cpu().numpy(), exact)
    taus = {str(t): choose_tau(sigma, err, t) for t in TARGETS}
    cal = {"weights": os.path.basename(a.weights), "n": a.n, "seed0": a.seed0, "tau": taus,
           "porosity_range": list(POROSITY_RANGE), "calibration_set": report(sigma, err, taus)}
    with open(a.out, "w") as f:
        json.dump(cal, f, indent=1)
    print(json.dumps(cal))
    return 0


if __name__ == "__main__":
    sys.exit(main())
