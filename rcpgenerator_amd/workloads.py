"""Synthetic packing workloads of the batched benchmarks.

``batch_cases`` draws packing i from its own seed (i): N uniform in
[nmin, nmax], a dimension from `dims`, a diameter family (mono;
bidisperse 1:1.4 at 50/50; log-normal with coefficient of variation
U[0.1, 0.3]; power law P(d) ~ d^-2.5 with size ratio U[2, 10]), periodic or
hard wall per axis and an initializer phi U[0.05, 0.15] (scaled by 0.4
from five dimensions up).  The defaults are
the batched configuration of 4096 packings of N 500..20000 in 2-4
dimensions.
"""
from __future__ import annotations

import math
import random


def batch_cases(P: int, seed0: int = 0, nmin: int = 500, nmax: int = 20000,
                dims=(2, 3, 4)) -> list:
    cases = []
    for i in range(seed0, seed0 + P):
        r = random.Random(i)
        N = r.randint(nmin, nmax)
        nd = r.choice(tuple(dims))
        fam = r.choice(("mono", "bidisperse", "lognormal", "powerlaw"))
        if fam == "mono":
            dist = {"type": "mono", "d": 1.0}
        elif fam == "bidisperse":
            dist = {"type": "bidisperse", "d1": 1.0, "d2": 1.4, "p": 0.5}
        elif fam == "lognormal":
            cv = r.uniform(0.1, 0.3)
            s2 = math.log(1.0 + cv * cv)
            dist = {"type": "lognormal", "mu": -0.5 * s2, "sigma": math.sqrt(s2)}
        else:
            ratio = r.uniform(2.0, 10.0)
            dist = {"type": "powerlaw", "d_min": 1.0 / ratio, "d_max": 1.0, "exponent": -2.5}
        walls = [r.choice((0, 1)) for _ in range(nd)]
        phi = r.uniform(0.05, 0.15)
        if nd >= 5:
            # random sequential addition saturates near 0.1 in five dimensions
            phi = 0.4 * phi
        cases.append({"phi": phi, "N": N, "Ndim": nd, "box": [1.0] * nd, "walls": walls,
                      "seed": i, "dist": dist})
    return cases
