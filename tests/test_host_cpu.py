"""Host-side checks that need no GPU: the C-ABI library exports every symbol
include/rcp_amd.h declares, the initializer (host C++) reproduces the
reference's initialize_particles, and the Packing API's host logic."""
from __future__ import annotations

import ctypes
import math
import os
import re

import numpy as np
import pytest

from conftest import ROOT, load_build_recipe, load_golden


@pytest.fixture(scope="session")
def native():
    _build = load_build_recipe()
    if not _build.built():
        _build.build()
    import rcpgenerator_amd
    return rcpgenerator_amd


def test_library_exports_every_declared_symbol(native):
    header = open(os.path.join(ROOT, "include", "rcp_amd.h")).read()
    declared = set(re.findall(r"\b(rcp_[a-z_0-9]+)\s*\(", header))
    declared -= {"rcp_observer_fn"}
    lib = ctypes.CDLL(native.native_library_path())
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    from rcpgenerator_amd import capi
    assert declared == set(capi.SYMBOLS)


def test_module_exports_reference_functions(native):
    for name in ("initialize_particles", "run_packing", "run_packing_observed",
                 "set_num_threads", "get_num_threads", "Packing"):
        assert hasattr(native, name)


def test_set_num_threads_contract(native):
    with pytest.raises(ValueError):
        native.set_num_threads(0)
    native.set_num_threads(3)
    assert native.get_num_threads() == 3
    native.set_num_threads(1)


def _init_cfg(name, nd, walls, dist):
    return {"phi": 0.1, "N": 40, "Ndim": nd, "box": [1.0] * nd, "walls": walls, "seed": 11,
            "dist": dist}


def test_initializer_matches_reference_golden(native):
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_golden", os.path.join(ROOT, "tests", "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    g = load_golden("init.npz")
    for name, dist, walls in mg.INIT_CASES:
        for nd in (2, 3):
            key = f"{name}_{nd}"
            w = [int(v) for v in g[key + "_walls"]]
            r = native.initialize_particles(_init_cfg(name, nd, w, dist))
            assert np.array_equal(np.array(r["diameters"]), g[key + "_D"]), key
            assert np.array_equal(np.array(r["positions"]), g[key + "_x"]), key
            assert np.array_equal(np.array(r["box"]), g[key + "_box"]), key
            assert r["diameter_scale_factor"] == float(g[key + "_factor"]), key
            assert r["phi_modifier"] == float(g[key + "_phimod"]), key


@pytest.mark.parametrize("N,nd,walls,dist,phi", [
    (2000, 3, [0, 0, 0], {"type": "bidisperse", "d1": 1.0, "d2": 1.4, "p": 0.5}, 0.08),
    (3000, 2, [0, 0], {"type": "lognormal", "mu": -0.5 * math.log(1.04),
                       "sigma": math.sqrt(math.log(1.04))}, 0.10),
    (3000, 3, [0, 0, 1], {"type": "powerlaw", "d_min": 0.1, "d_max": 1.0, "exponent": -3.0},
     0.05),
    (1000, 3, [-3, -1, -1], {"type": "powerlaw", "d_min": 0.1, "d_max": 1.0, "exponent": -3.0},
     0.05),
    (800, 4, [0, 0, 0, 0], {"type": "gaussian", "mu": 1.0, "sigma": 0.2}, 0.1),
    (300, 2, [0, 0], {"type": "mono", "d": 1.0}, 0.08),
])
def test_initializer_matches_reference(native, ref, N, nd, walls, dist, phi):
    mod = ref.ref_module()
    cfg = {"phi": phi, "N": N, "Ndim": nd, "box": [1.0] * nd, "walls": walls, "seed": 123,
           "dist": dist}
    a = mod.initialize_particles(cfg)
    b = native.initialize_particles(cfg)
    for k in ("positions", "diameters", "box", "walls", "diameter_scale_factor", "phi_modifier"):
        assert a[k] == b[k], k


def test_initializer_fix_height_and_errors(native, ref):
    mod = ref.ref_module()
    cfg = {"phi": 0.1, "N": 50, "Ndim": 2, "box": [1.0, 0.3], "walls": [0, 1],
           "fix_height": True, "seed": 9, "dist": {"type": "flat", "d_min": 0.5, "d_max": 1.0}}
    assert mod.initialize_particles(cfg) == native.initialize_particles(cfg)
    with pytest.raises(RuntimeError):
        native.initialize_particles({"phi": 0.1, "N": 10, "Ndim": 1})
    with pytest.raises(RuntimeError):
        native.initialize_particles({"phi": 1.5, "N": 10, "Ndim": 2})
    with pytest.raises(RuntimeError):
        native.initialize_particles({"phi": 0.1, "N": 10, "Ndim": 2, "dist": {"type": "nope"}})
    with pytest.raises(RuntimeError):   # RSA cannot place at this density
        native.initialize_particles({"phi": 0.9, "N": 400, "Ndim": 2})


def test_packing_host_logic(native):
    P = native.Packing
    with pytest.raises(TypeError, match="Unexpected keyword argument"):
        P(N=10, Ndim=2, box=[1, 1], bogus=1)
    p = P(phi=0.1, N=30, Ndim=2, box=[1.0, 1.0], seed=4)
    assert p.initialized and not p.packed and p.needs_pack and not p.needs_initialize
    assert len(p.positions) == 30 and len(p.diameters) == 30
    assert p.walls == [0, 0]
    assert p._current_phi() == pytest.approx(0.1, rel=1e-12)
    # dirty-state flags
    p.seed = 5
    assert p.needs_pack and not p.needs_initialize
    p.N = 31
    assert p.needs_initialize
    p.initialize()
    assert len(p.positions) == 31
    # update_phi / update_box keep their closed forms
    got = p.update_phi(0.2)
    assert got == pytest.approx(0.2, rel=1e-12)
    box = p.update_box([2.0, 1.0])
    assert box == [2.0, 1.0]
    assert p._current_phi() == pytest.approx(0.2, rel=1e-12)
    with pytest.raises(ValueError):
        p.update_box([1.0])
    with pytest.raises(ValueError):
        p.update_phi(-1)
    with pytest.raises(ValueError):
        p.relax(0)
    with pytest.raises(ValueError):
        p.relax(10, fix_diameter=True, target_phi=0.5)
    # serialisation round trip
    q = P.from_dict(p.to_dict())
    assert q.to_dict() == p.to_dict()
    assert repr(q).startswith("Packing(N=31")
    p.reset()
    assert not p.initialized and p.positions == []


def test_container_fraction():
    from rcpgenerator_amd._wrapper import _container_fraction
    assert _container_fraction([0, 0]) == 1.0
    assert _container_fraction([-2, -1]) == pytest.approx(math.pi / 4)
    assert _container_fraction([-3, -1, -1]) == pytest.approx(math.pi / 6)
    assert _container_fraction([-1, -1]) == 1.0
