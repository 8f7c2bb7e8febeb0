"""Writes tests/golden/scipy_fixtures.json: library known-answer values that pin
the oracle's special functions and statistics.

Run from the repository root: ``python tests/golden/make_golden.py``.
Sources: scipy.special.betainc, scipy.stats.t, scipy.stats.norm,
scipy.stats.ttest_ind_from_stats (Welch), scipy.stats.jarque_bera,
scipy.stats.skew / kurtosis (biased).  Every value is computed here, in this
container, from the scipy that is installed; no reference code is involved.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import scipy
from scipy import special, stats

OUT = Path(__file__).with_name("scipy_fixtures.json")


def main() -> None:
    rng = np.random.default_rng(20251203)
    fx: dict = {"scipy_version": scipy.__version__}

    # regularised incomplete beta, including the Student-t arguments (b = 1/2)
    cases = [(0.5, 0.5, 0.3), (2.0, 3.0, 0.4), (10.0, 0.5, 0.95), (31.5, 0.5, 0.8), (127.5, 0.5, 0.96),
             (2047.5, 0.5, 0.996), (4000.0, 0.5, 0.998), (1.0, 1.0, 0.25), (5.0, 0.5, 0.01), (0.5, 5.0, 0.99)]
    fx["betainc"] = [[a, b, x, float(special.betainc(a, b, x))] for a, b, x in cases]

    # two-sided t tail for t^2 and critical t^2 at alpha = 0.005
    nus = [1.0, 2.0, 3.5, 7.0, 15.0, 31.0, 63.0, 127.0, 255.0, 510.0, 1022.0, 4095.0, 8190.0]
    fx["t2_tail"] = []
    for nu in nus:
        for t2 in (0.5, 4.0, 7.88, 9.0, 25.0):
            fx["t2_tail"].append([t2, nu, float(2 * stats.t.sf(np.sqrt(t2), nu))])
    fx["t2_crit_0.005"] = [[nu, float(stats.t.ppf(1 - 0.0025, nu) ** 2)] for nu in nus]
    fx["z_0.9975"] = float(stats.norm.ppf(0.9975))

    # Welch tests from summary statistics (mean, sample sd, n)
    fx["welch"] = []
    for _ in range(24):
        n1, n2 = int(rng.integers(4, 600)), int(rng.integers(4, 600))
        m1, m2 = rng.normal(0, 1), rng.normal(0, 1)
        s1, s2 = rng.uniform(0.2, 5), rng.uniform(0.2, 5)
        m2 = m1 + rng.normal(0, 1) * np.sqrt(s1 * s1 / n1 + s2 * s2 / n2) * 3
        res = stats.ttest_ind_from_stats(m1, s1, n1, m2, s2, n2, equal_var=False)
        fx["welch"].append([m1, s1, n1, m2, s2, n2, float(res.statistic), float(res.pvalue)])

    # Jarque-Bera, skewness and kurtosis (biased) on raw samples
    fx["jb"] = []
    for k in range(6):
        n = [16, 64, 256, 1000, 37, 512][k]
        x = rng.lognormal(0, [0.3, 0.8, 1.2, 0.5, 1.0, 0.1][k], n).astype(np.float32)
        jb = stats.jarque_bera(x.astype(np.float64))
        fx["jb"].append({"x": x.tolist(), "jb": float(jb.statistic),
                         "skew": float(stats.skew(x.astype(np.float64), bias=True)),
                         "kurt": float(stats.kurtosis(x.astype(np.float64), fisher=False, bias=True))})
    OUT.write_text(json.dumps(fx, indent=1))
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
