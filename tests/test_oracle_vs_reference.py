"""Oracle vs the reference engine compiled from its own sources (oracle/_ref),
on seeded random inputs beyond the committed fixtures: bit-for-bit."""
from __future__ import annotations

import math

import numpy as np
import pytest


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _random(rng, N, nd, phi, poly=0.2):
    D = np.exp(rng.normal(0, poly, N))
    unit = math.pi ** (nd / 2) / math.gamma(nd / 2 + 1)
    D *= (phi / (unit * np.sum((D / 2) ** nd))) ** (1 / nd)
    return rng.random((N, nd)), D


CASES = [([0, 0], 0.8), ([1, 0], 0.7), ([-2, -1], 0.6), ([0, 0, 0], 0.55), ([1, 1, 1], 0.5),
         ([-3, -1, -1], 0.45), ([0, 0, 0, 0], 0.3), ([0, 0, 0, 0, 1], 0.2)]


@pytest.mark.parametrize("walls,phi", CASES)
def test_forces_bit_exact(ref, walls, phi):
    nd = len(walls)
    rng = np.random.default_rng(100 + nd + (walls[0] < 0))
    for trial in range(3):
        x, D = _random(rng, 150, nd, phi, poly=0.1 + 0.1 * trial)
        box = np.ones(nd)
        rows = ref.pairs_full(x, D, box)
        rows = (rows[0], rows[1])
        # reference rows must not repeat partners for this comparison
        off, nbr = rows
        uniq = [np.unique(nbr[off[i]:off[i + 1]]) for i in range(len(D))]
        off = np.concatenate([[0], np.cumsum([len(u) for u in uniq])]).astype(np.uint64)
        nbr = np.concatenate(uniq).astype(np.uint32)
        mu = 1e-3 * (trial + 1)
        Fr, sr = ref.ref_forces(x, D, box, walls, mu, (off, nbr))
        Fo, so = ref.forces(x, D, box, walls, mu)
        assert np.array_equal(_bits(Fr), _bits(Fo))
        for k in ("U", "max_min_dist", "Lc", "Fmean", "dkappa"):
            assert sr[k] == so[k], k


def test_forces_cascade_weights_bit_exact(ref):
    rng = np.random.default_rng(9)
    x, D = _random(rng, 120, 3, 0.5)
    fixed = rng.random(120) < 0.3
    rk = np.where(fixed, 0.0, D * 0.5)
    sum_active = float(np.sum(D[~fixed]))
    rows = ref.pairs_full(x, D, np.ones(3))
    Fr, sr = ref.ref_forces(x, D, np.ones(3), [0, 0, 1], 2e-3, rows, rk=rk,
                            sumD_active=sum_active)
    Fo, so = ref.forces(x, D, np.ones(3), [0, 0, 1], 2e-3, rk=rk, sumD_active=sum_active)
    assert np.array_equal(_bits(Fr), _bits(Fo))
    assert sr["dkappa"] == so["dkappa"]


@pytest.mark.parametrize("nd,backend", [(2, "kdtree"), (3, "kdtree"), (4, "sortwalk"),
                                        (5, "sortwalk")])
def test_pairs_bit_exact(ref, nd, backend):
    rng = np.random.default_rng(nd)
    for N, phi in ((200, 0.3), (400, 0.5 if nd <= 3 else 0.2)):
        x, D = _random(rng, N, nd, phi)
        a = ref.ref_pairs(x, D, np.ones(nd), backend=backend)
        b = ref.pairs_full(x, D, np.ones(nd))
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_mean_force_and_scale_bit_exact(ref):
    rng = np.random.default_rng(4)
    F = rng.normal(size=(333, 3)) * 1e-4
    F[::3] = 0.0
    assert ref.ref_shim().ref_mean_force(np.ascontiguousarray(F), 333, 3) == ref.mean_force(F)
    D = rng.random(77) + 0.2
    box = np.array([1.0, 1.2, 0.9])
    out_r = np.zeros(77)
    f_r = ref.ref_shim().ref_scale_diameters(D, 77, 0.33, box, 3, 0, out_r)
    out_o, f_o = ref.scale_diameters(D, 0.33, box, 3)
    assert f_r == f_o and np.array_equal(out_r, out_o)


RUNS = [
    (60, 2, [0, 0], {"type": "bidisperse", "d1": 1.0, "d2": 1.4, "p": 0.5}, {}),
    (60, 3, [0, 0, 0], {"type": "lognormal", "mu": 0.0, "sigma": 0.2}, {}),
    (50, 3, [1, 1, 1], {"type": "mono", "d": 1.0}, {"max_steps": 3000}),
    (60, 2, [-1, -1], {"type": "powerlaw", "d_min": 0.3, "d_max": 1.0, "exponent": -3.0}, {}),
    (40, 4, [0, 0, 0, 0], {"type": "mono", "d": 1.0}, {"max_steps": 4000}),
    (80, 2, [0, 0], {"type": "mono", "d": 1.0}, {"max_steps": 2500, "mu": 1e-3,
                                                 "target_phi": 0.7}),
    (80, 2, [0, 1], {"type": "flat", "d_min": 0.5, "d_max": 1.5},
     {"max_steps": 1200, "reorder_interval": 300, "initial_phi_target": 0.6}),
    (80, 2, [0, 0], {"type": "mono", "d": 1.0}, {"max_steps": 900, "fix_diameter": True}),
]


@pytest.mark.parametrize("N,nd,walls,dist,opts", RUNS)
def test_full_run_bit_exact(ref, N, nd, walls, dist, opts):
    mod = ref.ref_module()
    mod.set_num_threads(1)
    init = mod.initialize_particles({"phi": 0.1, "N": N, "Ndim": nd, "box": [1.0] * nd,
                                     "walls": walls, "seed": 5, "dist": dist})
    r = mod.run_packing_observed({"positions": init["positions"], "diameters": init["diameters"]},
                                 {"box": init["box"], "walls": init["walls"]},
                                 options=opts)["result"]
    o = ref.run_packing(init["positions"], init["diameters"], box=init["box"],
                        walls=init["walls"], max_steps=opts.get("max_steps", 0),
                        mu=opts.get("mu"), fix_diameter=opts.get("fix_diameter", False),
                        target_phi=opts.get("target_phi"),
                        reorder_interval=opts.get("reorder_interval", 2000),
                        initial_phi_target=opts.get("initial_phi_target"))
    assert r["steps"] == o["steps"]
    assert np.array_equal(_bits(np.array(r["positions"])), _bits(o["positions"]))
    assert np.array_equal(_bits(np.array(r["diameters"])), _bits(o["diameters"]))
    assert np.array_equal(np.array(r["particle_origin"], dtype=np.uint64), o["particle_origin"])
    for key in ("phi_history", "force_history", "energy_history", "max_overlap_history",
                "mu_history", "alpha_history"):
        assert np.array_equal(_bits(np.array(r[key])), _bits(o[key])), key
    assert r["phi"] == o["phi"] and r["max_min_dist"] == o["max_min_dist"]


@pytest.mark.parametrize("max_steps", [50, 300])
def test_fix_height_run_bit_exact(ref, max_steps):
    """fix_height (the box height follows kappa every step, RG:5773-5779)."""
    mod = ref.ref_module()
    mod.set_num_threads(1)
    init = mod.initialize_particles({"phi": 0.1, "N": 60, "Ndim": 2, "box": [1.0, 1.0],
                                     "walls": [0, 1], "seed": 9, "fix_height": True,
                                     "dist": {"type": "flat", "d_min": 0.5, "d_max": 1.0}})
    r = mod.run_packing_observed({"positions": init["positions"], "diameters": init["diameters"]},
                                 {"box": init["box"], "walls": init["walls"], "fix_height": True},
                                 options={"max_steps": max_steps})["result"]
    o = ref.run_packing(init["positions"], init["diameters"], box=init["box"],
                        walls=init["walls"], fix_height=True, max_steps=max_steps)
    assert r["steps"] == o["steps"]
    assert np.array_equal(_bits(np.array(r["positions"])), _bits(o["positions"]))
    assert np.array_equal(_bits(np.array(r["box"])), _bits(o["box"]))
    assert np.array_equal(_bits(np.array(r["phi_history"])), _bits(o["phi_history"]))


def test_cascade_run_bit_exact(ref):
    """The frozen-diameter mask (PackingInput.fixed, rcp_generator.hpp:30-32)."""
    mod = ref.ref_module()
    init = mod.initialize_particles({"phi": 0.1, "N": 60, "Ndim": 2, "box": [1.0, 1.0],
                                     "walls": [0, 0], "seed": 2,
                                     "dist": {"type": "bidisperse", "d1": 1.0, "d2": 2.0,
                                              "p": 0.5}})
    fixed = [1 if d > 0.9 * max(init["diameters"]) else 0 for d in init["diameters"]]
    r = mod.run_packing_observed({"positions": init["positions"], "diameters": init["diameters"],
                                  "fixed": fixed}, {"box": [1.0, 1.0], "walls": [0, 0]},
                                 options={"max_steps": 7000})["result"]
    o = ref.run_packing(init["positions"], init["diameters"], box=[1.0, 1.0], walls=[0, 0],
                        fixed=fixed, max_steps=7000)
    assert r["steps"] == o["steps"]
    assert np.array_equal(_bits(np.array(r["positions"])), _bits(o["positions"]))
    assert np.array_equal(_bits(np.array(r["force_history"])), _bits(o["force_history"]))
