"""ctypes front-end to the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/lib/liboracle.so`` (the C restatement in ``rcp_oracle.c``) and,
when it has been built, ``oracle/_ref`` (the reference engine compiled from its
own sources).  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this module, as the checker; the product package
never does.
"""
from __future__ import annotations

import ctypes as C
import importlib.machinery
import importlib.util
import os
import subprocess
import sysconfig

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
EXT_SUFFIX = sysconfig.get_config_var("EXT_SUFFIX") or ".so"

_lib = None
_shim = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


class ForceOut(C.Structure):
    _fields_ = [
        ("U", C.c_double), ("Lc", C.c_double), ("Fmean", C.c_double),
        ("dkappa", C.c_double), ("max_min_dist", C.c_double),
        ("U_raw", C.c_double), ("Lc_raw", C.c_double), ("Fmean_raw", C.c_double),
        ("dkappa_raw", C.c_double), ("sumD", C.c_double),
        ("count", C.c_uint64), ("pair_contacts", C.c_uint64),
    ]


class RunIn(C.Structure):
    _fields_ = [
        ("N", C.c_size_t), ("ndim", C.c_size_t),
        ("x", C.c_void_p), ("D", C.c_void_p), ("fixed", C.c_void_p),
        ("box", C.c_void_p), ("nbox", C.c_size_t),
        ("walls", C.c_void_p), ("nwalls", C.c_size_t),
        ("fix_height", C.c_int), ("max_steps", C.c_size_t),
        ("has_mu", C.c_int), ("mu", C.c_double),
        ("fix_diameter", C.c_int),
        ("has_target_phi", C.c_int), ("target_phi", C.c_double),
        ("reorder_interval", C.c_size_t),
        ("has_initial_phi", C.c_int), ("initial_phi", C.c_double),
    ]


class RunOut(C.Structure):
    _fields_ = [
        ("x", C.c_void_p), ("D", C.c_void_p), ("origin", C.c_void_p),
        ("box", C.c_void_p), ("walls", C.c_void_p), ("hist_cap", C.c_size_t),
        ("phi_hist", C.c_void_p), ("F_hist", C.c_void_p), ("U_hist", C.c_void_p),
        ("maxov_hist", C.c_void_p), ("mu_hist", C.c_void_p), ("alpha_hist", C.c_void_p),
        ("muflag_hist", C.c_void_p),
        ("steps", C.c_size_t), ("phi", C.c_double), ("max_min_dist", C.c_double),
        ("force_magnitude", C.c_double), ("kappa", C.c_double), ("error", C.c_int),
    ]


def build() -> None:
    """Compile the C restatement (and the reference checker when the reference
    checkout is present)."""
    targets = ["oracle"]
    if os.path.isdir(os.environ.get("RCP_REFERENCE_SRC", "/root/reference/python_code/python")):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE, *targets], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_forces.argtypes = [C.c_size_t, C.c_size_t, _dp, _dp, _dp, _i8p, C.c_double,
                                 C.c_void_p, C.c_double, _dp, C.POINTER(ForceOut)]
        L.orc_forces_list.argtypes = [C.c_size_t, C.c_size_t, _dp, _dp, _dp, _i8p, C.c_double,
                                      _u64p, _u32p, C.c_void_p, C.c_double, _dp,
                                      C.POINTER(ForceOut)]
        L.orc_mean_force.restype = C.c_double
        L.orc_mean_force.argtypes = [_dp, C.c_size_t, C.c_size_t]
        L.orc_shell.restype = C.c_double
        L.orc_shell.argtypes = [_dp, C.c_size_t]
        L.orc_pairs_full.restype = C.c_size_t
        L.orc_pairs_full.argtypes = [C.c_size_t, C.c_size_t, _dp, _dp, _dp, _u64p,
                                     C.c_void_p, C.c_size_t]
        L.orc_reorder_perm.argtypes = [_dp, C.c_size_t, C.c_size_t, _dp, _u32p]
        L.orc_adam_positions.argtypes = [C.c_size_t, C.c_size_t, _dp, _dp, _dp, _dp, _dp,
                                         C.c_double, C.c_double, C.c_double]
        L.orc_scale_diameters.restype = C.c_double
        L.orc_scale_diameters.argtypes = [_dp, C.c_size_t, C.c_double, _dp, C.c_size_t,
                                          C.c_int, _dp]
        L.orc_step_cap.restype = C.c_size_t
        L.orc_step_cap.argtypes = [C.c_size_t, C.c_size_t, C.c_size_t]
        L.orc_run_packing.restype = C.c_int
        L.orc_run_packing.argtypes = [C.POINTER(RunIn), C.POINTER(RunOut)]
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def forces(x, D, box, walls, mu, rk=None, sumD_active=-1.0, rows=None):
    """One force evaluation.  Returns (F[N,ndim], dict of scalars)."""
    x = _f64(x)
    N, nd = x.shape
    D, box = _f64(D), _f64(box)
    walls = np.ascontiguousarray(walls, dtype=np.int8)
    F = np.zeros((N, nd))
    o = ForceOut()
    rkp = None if rk is None else _f64(rk)
    rk_ptr = None if rkp is None else rkp.ctypes.data
    if rows is None:
        lib().orc_forces(N, nd, x, D, box, walls, mu, rk_ptr, sumD_active, F, C.byref(o))
    else:
        off, nbr = rows
        lib().orc_forces_list(N, nd, x, D, box, walls, mu,
                              np.ascontiguousarray(off, dtype=np.uint64),
                              np.ascontiguousarray(nbr, dtype=np.uint32),
                              rk_ptr, sumD_active, F, C.byref(o))
    return F, {k: getattr(o, k) for k, _ in ForceOut._fields_}


def mean_force(F):
    F = _f64(F)
    return lib().orc_mean_force(F, F.shape[0], F.shape[1])


def shell(D):
    D = _f64(D)
    return lib().orc_shell(D, D.shape[0])


def pairs_full(x, D, box):
    """Full-refresh pair set of the reference's default backend, as CSR."""
    x, D, box = _f64(x), _f64(D), _f64(box)
    N, nd = x.shape
    off = np.zeros(N + 1, dtype=np.uint64)
    n = lib().orc_pairs_full(N, nd, x, D, box, off, None, 0)
    nbr = np.zeros(max(n, 1), dtype=np.uint32)
    lib().orc_pairs_full(N, nd, x, D, box, off, nbr.ctypes.data, n)
    return off, nbr[:n]


def reorder_perm(x, box):
    x, box = _f64(x), _f64(box)
    perm = np.zeros(x.shape[0], dtype=np.uint32)
    lib().orc_reorder_perm(x, x.shape[0], x.shape[1], box, perm)
    return perm


def adam_positions(F, m, v, x, box, alpha, inv_b1, inv_b2):
    F, m, v, x, box = _f64(F), _f64(m).copy(), _f64(v).copy(), _f64(x).copy(), _f64(box)
    lib().orc_adam_positions(x.shape[0], x.shape[1], F, m, v, x, box, alpha, inv_b1, inv_b2)
    return m, v, x


def scale_diameters(D, phi_target, box, ndim, fix_height=False):
    D, box = _f64(D), _f64(box)
    out = np.zeros_like(D)
    f = lib().orc_scale_diameters(D, D.shape[0], phi_target, box, ndim, int(fix_height), out)
    return out, f


def run_packing(positions, diameters, box=None, walls=None, fixed=None, fix_height=False,
                max_steps=0, mu=None, fix_diameter=False, target_phi=None,
                reorder_interval=2000, initial_phi_target=None):
    """Full run restated from run_packing_observed (default environment).
    Returns a dict with the reference result's keys (arrays, not lists)."""
    x = _f64(positions)
    N, nd = x.shape
    D = _f64(diameters)
    bx = _f64(box if box is not None else [])
    wl = np.ascontiguousarray(walls if walls is not None else [], dtype=np.int8)
    fx = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.int8)
    cap = lib().orc_step_cap(N, nd, max_steps) + 1
    ri = RunIn(N=N, ndim=nd, x=x.ctypes.data, D=D.ctypes.data,
               fixed=None if fx is None else fx.ctypes.data,
               box=bx.ctypes.data if bx.size else None, nbox=bx.size,
               walls=wl.ctypes.data if wl.size else None, nwalls=wl.size,
               fix_height=int(fix_height), max_steps=max_steps,
               has_mu=int(mu is not None), mu=float(mu or 0.0),
               fix_diameter=int(fix_diameter),
               has_target_phi=int(target_phi is not None), target_phi=float(target_phi or 0.0),
               reorder_interval=reorder_interval,
               has_initial_phi=int(initial_phi_target is not None),
               initial_phi=float(initial_phi_target or 0.0))
    ox = np.zeros((N, nd)); oD = np.zeros(N); og = np.zeros(N, dtype=np.uint64)
    ob = np.zeros(nd); ow = np.zeros(nd, dtype=np.int8)
    H = {k: np.zeros(cap) for k in ("phi", "F", "U", "ov", "mu", "al")}
    fl = np.zeros(cap, dtype=np.int8)
    ro = RunOut(x=ox.ctypes.data, D=oD.ctypes.data, origin=og.ctypes.data, box=ob.ctypes.data,
                walls=ow.ctypes.data, hist_cap=cap,
                phi_hist=H["phi"].ctypes.data, F_hist=H["F"].ctypes.data,
                U_hist=H["U"].ctypes.data, maxov_hist=H["ov"].ctypes.data,
                mu_hist=H["mu"].ctypes.data, alpha_hist=H["al"].ctypes.data,
                muflag_hist=fl.ctypes.data)
    rc = lib().orc_run_packing(C.byref(ri), C.byref(ro))
    if rc == 1:
        raise RuntimeError("Error: --box length mismatch")
    if rc == 2:
        raise RuntimeError("Error: --walls expects Ndim entries")
    s = ro.steps
    return {
        "positions": ox, "diameters": oD, "box": ob, "walls": ow,
        "phi_history": H["phi"][:s], "force_history": H["F"][:s],
        "energy_history": H["U"][:s], "max_overlap_history": H["ov"][:s],
        "mu_history": H["mu"][:s], "alpha_history": H["al"][:s],
        "mu_flag_history": fl[:s], "steps": s, "phi": ro.phi,
        "max_min_dist": ro.max_min_dist, "force_magnitude": ro.force_magnitude,
        "particle_origin": og, "kappa": ro.kappa,
    }


# ------------------------------------------------------------------ reference

def have_ref() -> bool:
    return os.path.exists(os.path.join(REF_DIR, "libref_shim.so"))


def ref_shim():
    global _shim
    if _shim is None:
        S = C.CDLL(os.path.join(REF_DIR, "libref_shim.so"))
        S.ref_set_threads.argtypes = [C.c_int]
        S.ref_forces.restype = C.c_int
        S.ref_forces.argtypes = [C.c_size_t, C.c_size_t, _dp, _dp, _dp, _i8p, C.c_double,
                                 _u64p, _u32p, C.c_void_p, C.c_double, _dp, _dp]
        S.ref_pairs.restype = C.c_longlong
        S.ref_pairs.argtypes = [C.c_int, C.c_size_t, C.c_size_t, _dp, _dp, _dp, C.c_uint32,
                                _u64p, C.c_void_p, C.c_size_t]
        S.ref_mean_force.restype = C.c_double
        S.ref_mean_force.argtypes = [_dp, C.c_size_t, C.c_size_t]
        S.ref_scale_diameters.restype = C.c_double
        S.ref_scale_diameters.argtypes = [_dp, C.c_size_t, C.c_double, _dp, C.c_size_t,
                                          C.c_int, _dp]
        S.ref_set_threads(1)
        _shim = S
    return _shim


def ref_forces(x, D, box, walls, mu, rows, rk=None, sumD_active=-1.0):
    x, D, box = _f64(x), _f64(D), _f64(box)
    N, nd = x.shape
    walls = np.ascontiguousarray(walls, dtype=np.int8)
    off, nbr = rows
    F = np.zeros((N, nd))
    out = np.zeros(8)
    rkp = None if rk is None else _f64(rk)
    rc = ref_shim().ref_forces(N, nd, x, D, box, walls, mu,
                               np.ascontiguousarray(off, dtype=np.uint64),
                               np.ascontiguousarray(nbr if len(nbr) else [0], dtype=np.uint32),
                               None if rkp is None else rkp.ctypes.data, sumD_active, F, out)
    assert rc == 0
    return F, {"U": out[0], "max_min_dist": out[1], "Lc": out[2], "Fmean": out[3],
               "dkappa": out[4]}


def ref_pairs(x, D, box, backend="kdtree", cap_per_row=None):
    x, D, box = _f64(x), _f64(D), _f64(box)
    N, nd = x.shape
    cap = cap_per_row or N
    off = np.zeros(N + 1, dtype=np.uint64)
    be = 0 if backend == "kdtree" else 1
    n = ref_shim().ref_pairs(be, N, nd, x, D, box, cap, off, None, 0)
    if n < 0:
        raise RuntimeError("reference pair build failed")
    nbr = np.zeros(max(n, 1), dtype=np.uint32)
    ref_shim().ref_pairs(be, N, nd, x, D, box, cap, off, nbr.ctypes.data, n)
    return off, nbr[:n]


def ref_module(native: bool = False):
    """The reference's own pybind module (_rcpgenerator), loaded from
    oracle/_ref without entering sys.modules."""
    path = os.path.join(REF_DIR, "native" if native else "", "_rcpgenerator" + EXT_SUFFIX)
    loader = importlib.machinery.ExtensionFileLoader("_rcpgenerator", path)
    spec = importlib.util.spec_from_file_location("_rcpgenerator", path, loader=loader)
    mod = importlib.util.module_from_spec(spec)
    loader.exec_module(mod)
    return mod
