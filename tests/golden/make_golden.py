"""Writes tests/golden/*.npz from the reference engine built by oracle/Makefile.

The reference (KD-physics/RCPGenerator, C++ core) ships no golden vectors,
so these fixtures are produced by its own code, compiled from its sources
into oracle/_ref, on small seeded inputs:

  forces.npz   get_forces_nd_3 (rcp_generator.cpp:3597-4269) on random
               configurations: 2-5 D, periodic / flat walls / curved.
  pairs.npz    get_pairs_kdtree (:2758) and get_pairs_nd_3 (:1198) on full
               refresh.
  runs.npz     run_packing (:4337) on the acceptance cases of
               tests/packer_parity.cpp:490-500 and tests/wheel_smoke.py:12-20
               and a hard-wall case (tests/hard_wall_regression.cpp:36-82):
               final state, step counts and SHA-256 digests of every history.
  init.npz     initialize_particles (initialize_particles.cpp:399-474) for
               every distribution type.

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

FORCE_CASES = [
    # (ndim, walls, N, phi, polydispersity)
    (2, [0, 0], 60, 0.80, 0.2),
    (2, [1, 1], 60, 0.70, 0.2),
    (2, [-2, -1], 50, 0.60, 0.1),
    (3, [0, 0, 0], 60, 0.55, 0.2),
    (3, [0, 0, 1], 60, 0.50, 0.3),
    (3, [-3, -1, -1], 50, 0.40, 0.1),
    (3, [-2, -1, 0], 50, 0.45, 0.1),
    (4, [0, 0, 0, 0], 60, 0.30, 0.2),
    (5, [0, 0, 0, 1, 1], 60, 0.20, 0.2),
]


def random_config(rng, N, nd, phi, poly, walls):
    D = np.exp(rng.normal(0.0, poly, N))
    unit = math.pi ** (nd / 2) / math.gamma(nd / 2 + 1)
    D *= (phi / (unit * np.sum((D / 2) ** nd))) ** (1.0 / nd)
    x = rng.random((N, nd))
    if walls[0] < 0:
        t = -walls[0]
        x[:, :t] = 0.5 + (x[:, :t] - 0.5) * 0.8
    return x, D, np.ones(nd)


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def make_forces(rng):
    out = {}
    for k, (nd, walls, N, phi, poly) in enumerate(FORCE_CASES):
        x, D, box = random_config(rng, N, nd, phi, poly, walls)
        off, nbr = O.pairs_full(x, D, box)
        mu = 2.5e-4 * (1 + k)
        F, s = O.ref_forces(x, D, box, walls, mu, (off, nbr))
        out[f"c{k}_x"] = x
        out[f"c{k}_D"] = D
        out[f"c{k}_box"] = box
        out[f"c{k}_walls"] = np.array(walls, dtype=np.int8)
        out[f"c{k}_mu"] = np.array(mu)
        out[f"c{k}_F"] = F
        out[f"c{k}_scalars"] = np.array([s["U"], s["max_min_dist"], s["Lc"], s["Fmean"],
                                         s["dkappa"]])
    out["ncases"] = np.array(len(FORCE_CASES))
    np.savez_compressed(os.path.join(HERE, "forces.npz"), **out)


def make_pairs(rng):
    out = {}
    cases = [(2, 300, 0.5), (3, 300, 0.4), (4, 200, 0.25), (5, 200, 0.15)]
    for k, (nd, N, phi) in enumerate(cases):
        x, D, box = random_config(rng, N, nd, phi, 0.2, [0] * nd)
        backend = "kdtree" if nd <= 3 else "sortwalk"
        off, nbr = O.ref_pairs(x, D, box, backend=backend)
        out[f"p{k}_x"] = x
        out[f"p{k}_D"] = D
        out[f"p{k}_off"] = off
        out[f"p{k}_nbr"] = nbr
    out["ncases"] = np.array(len(cases))
    np.savez_compressed(os.path.join(HERE, "pairs.npz"), **out)


RUN_CASES = {
    # tests/packer_parity.cpp:490-500
    "four_disks": dict(positions=[[.10, .10], [.35, .25], [.70, .55], [.82, .85]],
                       diameters=[.12, .10, .09, .08], box=[1.0, 1.0], walls=[0, 0]),
}


def make_runs():
    ref = O.ref_module()
    ref.set_num_threads(1)
    cases = dict(RUN_CASES)
    # tests/wheel_smoke.py:12-20
    init = ref.initialize_particles({"phi": 0.08, "N": 16, "Ndim": 2, "box": [1.0, 1.0],
                                     "walls": [0, 0], "dist": {"type": "mono", "d": 1.0},
                                     "seed": 20240719})
    cases["wheel_smoke"] = dict(positions=init["positions"], diameters=init["diameters"],
                                box=init["box"], walls=init["walls"])
    # tests/hard_wall_regression.cpp:36-82 (N=64 mono 2-D, walls {1,1}, 500 steps)
    init = ref.initialize_particles({"phi": 0.11, "N": 64, "Ndim": 2, "box": [1.0, 1.0],
                                     "walls": [1, 1], "dist": {"type": "mono", "d": 1.0},
                                     "seed": 1})
    cases["hard_wall"] = dict(positions=init["positions"], diameters=init["diameters"],
                              box=init["box"], walls=init["walls"], max_steps=500)
    # a 3-D bidisperse case (config-1 distribution at N=64) and a curved container
    init = ref.initialize_particles({"phi": 0.08, "N": 64, "Ndim": 3, "box": [1.0] * 3,
                                     "walls": [0, 0, 0], "seed": 123,
                                     "dist": {"type": "bidisperse", "d1": 1.0, "d2": 1.4,
                                              "p": 0.5}})
    cases["bidisperse_3d"] = dict(positions=init["positions"], diameters=init["diameters"],
                                  box=init["box"], walls=init["walls"])
    init = ref.initialize_particles({"phi": 0.1, "N": 48, "Ndim": 2, "box": [1.0, 1.0],
                                     "walls": [-1, -1], "seed": 3,
                                     "dist": {"type": "lognormal", "mu": 0.0, "sigma": 0.2}})
    cases["disk_container"] = dict(positions=init["positions"], diameters=init["diameters"],
                                   box=init["box"], walls=init["walls"])
    out = {"names": np.array(sorted(cases))}
    for name in sorted(cases):
        c = cases[name]
        opts = {"max_steps": c["max_steps"]} if "max_steps" in c else {}
        r = ref.run_packing_observed({"positions": c["positions"], "diameters": c["diameters"]},
                                     {"box": c["box"], "walls": c["walls"]},
                                     options=opts)["result"]
        out[f"{name}_x"] = np.array(c["positions"], dtype=np.float64)
        out[f"{name}_D"] = np.array(c["diameters"], dtype=np.float64)
        out[f"{name}_box"] = np.array(c["box"], dtype=np.float64)
        out[f"{name}_walls"] = np.array(c["walls"], dtype=np.int8)
        out[f"{name}_max_steps"] = np.array(c.get("max_steps", 0))
        out[f"{name}_steps"] = np.array(r["steps"])
        out[f"{name}_phi"] = np.array(r["phi"])
        out[f"{name}_final_x"] = np.array(r["positions"], dtype=np.float64)
        out[f"{name}_final_D"] = np.array(r["diameters"], dtype=np.float64)
        out[f"{name}_origin"] = np.array(r["particle_origin"], dtype=np.uint64)
        for key in ("phi_history", "force_history", "energy_history", "max_overlap_history",
                    "mu_history", "alpha_history"):
            out[f"{name}_{key}_sha256"] = np.array(digest(np.array(r[key], dtype=np.float64)))
        out[f"{name}_mu_flag_history_sha256"] = np.array(
            digest(np.array(r["mu_flag_history"], dtype=np.int8)))
        out[f"{name}_phi_head"] = np.array(r["phi_history"][:64], dtype=np.float64)
        out[f"{name}_F_head"] = np.array(r["force_history"][:64], dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, "runs.npz"), **out)


INIT_CASES = [
    ("mono", {"type": "mono", "d": 1.0}, [0, 0]),
    ("gaussian", {"type": "gaussian", "mu": 1.0, "sigma": 0.2}, [0, 0]),
    ("bigaussian", {"type": "bigaussian", "mu1": 1.0, "sigma1": 0.1, "mu2": 2.0,
                    "sigma2": 0.2, "p": 0.4}, [1, 0]),
    ("bidisperse", {"type": "bidisperse", "d1": 1.0, "d2": 1.4, "p": 0.5}, [0, 1]),
    ("lognormal", {"type": "lognormal", "mu": 0.0, "sigma": 0.3}, [0, 0]),
    ("flat", {"type": "flat", "d_min": 0.5, "d_max": 1.5}, [0, 0]),
    ("powerlaw", {"type": "powerlaw", "d_min": 0.1, "d_max": 1.0, "exponent": -3.0}, [0, 0]),
    ("exponential", {"type": "exponential", "d_min": 0.5, "d_max": 2.0}, [0, 0]),
    ("weibull", {"type": "weibull", "shape": 2.0, "scale": 1.0}, [-1, -1]),
    ("custom", {"type": "custom", "custom": [1.0 + 0.01 * k for k in range(40)]}, [0, 0]),
]


def make_init():
    ref = O.ref_module()
    out = {"names": np.array([c[0] for c in INIT_CASES])}
    for name, dist, walls in INIT_CASES:
        for nd in (2, 3):
            w = list(walls) + [0] * (nd - 2)
            cfg = {"phi": 0.1, "N": 40, "Ndim": nd, "box": [1.0] * nd, "walls": w,
                   "seed": 11, "dist": dist}
            r = ref.initialize_particles(cfg)
            key = f"{name}_{nd}"
            out[key + "_walls"] = np.array(w, dtype=np.int8)
            out[key + "_x"] = np.array(r["positions"], dtype=np.float64)
            out[key + "_D"] = np.array(r["diameters"], dtype=np.float64)
            out[key + "_box"] = np.array(r["box"], dtype=np.float64)
            out[key + "_factor"] = np.array(r["diameter_scale_factor"])
            out[key + "_phimod"] = np.array(r["phi_modifier"])
    np.savez_compressed(os.path.join(HERE, "init.npz"), **out)


def main():
    if not O.have_ref():
        O.build()
    rng = np.random.default_rng(20240719)
    make_forces(rng)
    make_pairs(rng)
    make_runs()
    make_init()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
