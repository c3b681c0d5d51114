"""Pins the CPU oracle against the golden vectors written by the reference
build (tests/golden/make_golden.py) -- bit-for-bit."""
from __future__ import annotations

import hashlib

import numpy as np

from conftest import load_golden


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def test_forces_match_reference_golden(oracle):
    g = load_golden("forces.npz")
    for k in range(int(g["ncases"])):
        x, D, box, walls, mu = (g[f"c{k}_{n}"] for n in ("x", "D", "box", "walls", "mu"))
        F, s = oracle.forces(x, D, box, walls, float(mu))
        assert np.array_equal(_bits(F), _bits(g[f"c{k}_F"])), f"case {k}: F differs"
        got = np.array([s["U"], s["max_min_dist"], s["Lc"], s["Fmean"], s["dkappa"]])
        assert np.array_equal(_bits(got), _bits(g[f"c{k}_scalars"])), f"case {k}: scalars"
        # the same evaluation on the explicit shell-inflated pair list (the
        # sort+walk list can repeat a partner in tiny boxes; keep one copy)
        off, nbr = oracle.pairs_full(x, D, box)
        rows_u = [np.unique(nbr[off[i]:off[i + 1]]) for i in range(len(D))]
        off_u = np.concatenate([[0], np.cumsum([len(r) for r in rows_u])]).astype(np.uint64)
        nbr_u = np.concatenate(rows_u).astype(np.uint32)
        F2, s2 = oracle.forces(x, D, box, walls, float(mu), rows=(off_u, nbr_u))
        assert np.array_equal(_bits(F2), _bits(F))
        assert s2["U"] == s["U"] and s2["dkappa"] == s["dkappa"]


def test_pairs_match_reference_golden(oracle):
    g = load_golden("pairs.npz")
    for k in range(int(g["ncases"])):
        x, D = g[f"p{k}_x"], g[f"p{k}_D"]
        off, nbr = oracle.pairs_full(x, D, np.ones(x.shape[1]))
        assert np.array_equal(off, g[f"p{k}_off"]), f"case {k}: row offsets"
        assert np.array_equal(nbr, g[f"p{k}_nbr"]), f"case {k}: partners"


def _digest(a, dtype=np.float64):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=dtype).tobytes()).hexdigest()


def test_full_runs_match_reference_golden(oracle):
    g = load_golden("runs.npz")
    for name in [str(n) for n in g["names"]]:
        r = oracle.run_packing(g[f"{name}_x"], g[f"{name}_D"], box=g[f"{name}_box"],
                               walls=g[f"{name}_walls"], max_steps=int(g[f"{name}_max_steps"]))
        assert r["steps"] == int(g[f"{name}_steps"]), name
        assert r["phi"] == float(g[f"{name}_phi"]), name
        assert np.array_equal(_bits(r["positions"]), _bits(g[f"{name}_final_x"])), name
        assert np.array_equal(_bits(r["diameters"]), _bits(g[f"{name}_final_D"])), name
        assert np.array_equal(r["particle_origin"], g[f"{name}_origin"]), name
        for key in ("phi_history", "force_history", "energy_history", "max_overlap_history",
                    "mu_history", "alpha_history"):
            assert _digest(r[key]) == str(g[f"{name}_{key}_sha256"]), (name, key)
        assert _digest(r["mu_flag_history"], np.int8) == str(g[f"{name}_mu_flag_history_sha256"])


def test_acceptance_four_disks(oracle):
    """tests/packer_parity.cpp:569-635 acceptance gates on the 4-disk case."""
    x = np.array([[.10, .10], [.35, .25], [.70, .55], [.82, .85]])
    D = np.array([.12, .10, .09, .08])
    a = oracle.run_packing(x, D, box=[1.0, 1.0], walls=[0, 0])
    b = oracle.run_packing(x, D, box=[1.0, 1.0], walls=[0, 0])
    assert np.array_equal(a["positions"], b["positions"])       # bitwise repeatable
    assert 0.75 < a["phi"] < 0.85
    assert a["max_min_dist"] < 5e-4
    assert a["steps"] < 60000
    # recomputed max overlap from the final state
    F, s = oracle.forces(a["positions"], a["diameters"], np.ones(2), np.zeros(2), 0.0)
    assert s["max_min_dist"] < 5e-4
    for key in ("phi_history", "force_history", "energy_history"):
        assert len(a[key]) == a["steps"]
