"""Stateful ``Packing`` API over the native ``_rcpgenerator`` module.

Behaviour follows the reference's public API (KD-physics/RCPGenerator,
python_code/python/README.md:84-533 and python/rcpgenerator/_wrapper.py):
constructor keywords, dirty-state flags, ``pack``/``relax``/``update_*``,
the result dictionary and the ``to_dict``/``from_dict`` schema.  Reported
packing fractions are referenced to the curved container's volume when
``walls[0] < 0`` (the reference's container correction, _wrapper.py:15-23,
309-332, including its second division of ``phi_history``).
"""
from __future__ import annotations

import math
from copy import deepcopy
from pathlib import Path
from typing import Any, Callable

from . import _rcpgenerator as _native

_DEFAULT_DIST = {"type": "mono", "d": 1.0}

# Attributes whose assignment invalidates the realized state / the result.
_INIT_FIELDS = frozenset({"phi", "N", "Ndim", "box", "walls", "fix_height", "dist"})
_PACK_FIELDS = frozenset({"box", "walls", "neighbor_max", "seed", "fix_height"})
_STATE_FIELDS = frozenset({"positions", "diameters"})

_RESULT_FIELDS = (
    "phi_history", "force_history", "energy_history", "max_overlap_history",
    "mu_flag_history", "mu_history", "alpha_history",
)
_TRAJ_FIELDS = ("positions", "diameters", "steps", "phi", "force", "energy", "max_min_dist")


def _ball_fraction(t: int) -> float:
    """Volume of a t-ball of unit diameter: the fraction of its bounding box."""
    return math.pi ** (t / 2.0) / math.gamma(t / 2.0 + 1.0) * 0.5 ** t


def _container_fraction(walls) -> float:
    """Box fraction filled by a curved container (walls[0] = -t, t >= 2)."""
    t = -int(walls[0]) if walls and int(walls[0]) < 0 else 0
    return _ball_fraction(t) if t >= 2 else 1.0


class Packing:
    """A packing: configuration, realized state and the last result."""

    _INIT_KEYS = set(_INIT_FIELDS)
    _PACK_KEYS = set(_PACK_FIELDS)
    _REALIZED_STATE_KEYS = set(_STATE_FIELDS)
    _RESULT_KEYS = {
        "steps", "phi_final", "max_min_dist", "force_magnitude",
        "phi_history", "force_history", "energy_history",
    }

    # ------------------------------------------------------------ lifecycle
    def __init__(self, **kwargs: Any) -> None:
        with self._internal():
            self._set_flags(initialized=False, packed=False, needs_initialize=True,
                            needs_pack=True)
            self.verbose = kwargs.pop("verbose", False)
            self.phi = kwargs.pop("phi", 0.05)
            self.N = kwargs.pop("N", 0)
            self.Ndim = kwargs.pop("Ndim", 0)
            self.box = deepcopy(kwargs.pop("box", []))
            self.walls = deepcopy(kwargs.pop("walls", [0] * len(self.box) if self.box else []))
            self.fix_height = kwargs.pop("fix_height", False)
            self.dist = deepcopy(kwargs.pop("dist", dict(_DEFAULT_DIST)))
            self.neighbor_max = kwargs.pop("neighbor_max", 0)
            self.seed = kwargs.pop("seed", 0)
            self.positions = []
            self.diameters = []
            self._reset_outputs()
            if kwargs:
                raise TypeError("Unexpected keyword argument(s): " + ", ".join(sorted(kwargs)))
        self.initialize()

    @classmethod
    def from_dict(cls, data: dict[str, Any]) -> "Packing":
        init = deepcopy(data.get("init_config", {}))
        pack = deepcopy(data.get("pack_config", {}))
        state = deepcopy(data.get("realized_state", {}))
        final = deepcopy(data.get("final_result", {}))
        traj = deepcopy(data.get("trajectory", {}))
        status = deepcopy(data.get("status", {}))
        obj = cls.__new__(cls)
        with obj._internal():
            obj._set_flags(initialized=False, packed=False, needs_initialize=True,
                           needs_pack=True)
            obj.verbose = init.pop("verbose", pack.pop("verbose", False))
            obj.phi = init.get("phi", 0.05)
            obj.N = init.get("N", 0)
            obj.Ndim = init.get("Ndim", 0)
            obj.box = deepcopy(init.get("box", pack.get("box", [])))
            obj.walls = deepcopy(init.get("walls", pack.get("walls", [])))
            obj.fix_height = init.get("fix_height", pack.get("fix_height", False))
            obj.dist = deepcopy(init.get("dist", {}))
            obj.neighbor_max = pack.get("neighbor_max", 0)
            obj.seed = pack.get("seed", 0)
            obj.positions = deepcopy(state.get("positions", []))
            obj.diameters = deepcopy(state.get("diameters", []))
            obj._reset_outputs()
            obj.steps = final.get("steps", 0)
            obj.phi_final = final.get("phi_final")
            obj.max_min_dist = final.get("max_min_dist")
            obj.force_magnitude = final.get("force_magnitude")
            for key in ("phi_history", "force_history", "energy_history"):
                setattr(obj, key, deepcopy(final.get(key, [])))
            for key in _TRAJ_FIELDS:
                setattr(obj, "trajectory_" + key, deepcopy(traj.get(key, [])))
        initialized = bool(status.get("initialized", bool(obj.positions) and bool(obj.diameters)))
        packed = bool(status.get("packed", bool(obj.steps)))
        obj._set_flags(
            initialized=initialized,
            packed=packed,
            needs_initialize=bool(status.get("needs_initialize", not initialized)),
            needs_pack=bool(status.get("needs_pack", not packed)),
        )
        return obj

    # ------------------------------------------------------ state plumbing
    class _Internal:
        def __init__(self, owner):
            self.owner = owner

        def __enter__(self):
            object.__setattr__(self.owner, "_mutating_internal_state", True)

        def __exit__(self, *exc):
            object.__setattr__(self.owner, "_mutating_internal_state", False)
            return False

    def _internal(self):
        return Packing._Internal(self)

    def _set_flags(self, **flags: bool) -> None:
        for name, value in flags.items():
            object.__setattr__(self, "_" + name, value)

    def __setattr__(self, name: str, value: Any) -> None:
        object.__setattr__(self, name, value)
        if name.startswith("_") or getattr(self, "_mutating_internal_state", False):
            return
        if name in _INIT_FIELDS:
            self._set_flags(needs_initialize=True, needs_pack=True, packed=False)
        elif name in _PACK_FIELDS or name in _STATE_FIELDS:
            self._set_flags(needs_pack=True, packed=False)

    def _reset_outputs(self) -> None:
        self.steps = 0
        self.phi_final = None
        self.max_min_dist = None
        self.force_magnitude = None
        for key in _RESULT_FIELDS:
            setattr(self, key, [])
        for key in _TRAJ_FIELDS:
            setattr(self, "trajectory_" + key, [])

    def _clear_results(self) -> None:
        self.steps = 0
        self.phi_final = None
        self.max_min_dist = None
        self.force_magnitude = None
        for key in _RESULT_FIELDS:
            setattr(self, key, [])

    def _clear_trajectory(self) -> None:
        for key in _TRAJ_FIELDS:
            setattr(self, "trajectory_" + key, [])

    @property
    def needs_initialize(self) -> bool:
        return self._needs_initialize

    @property
    def needs_pack(self) -> bool:
        return self._needs_pack

    @property
    def initialized(self) -> bool:
        return self._initialized

    @property
    def packed(self) -> bool:
        return self._packed

    def __repr__(self) -> str:
        return (f"Packing(N={self.N}, Ndim={self.Ndim}, initialized={self._initialized}, "
                f"packed={self._packed}, needs_initialize={self._needs_initialize}, "
                f"needs_pack={self._needs_pack})")

    def _log(self, message: str) -> None:
        if self.verbose:
            print(message)

    # ------------------------------------------------------------ payloads
    def _collect_init_config(self) -> dict[str, Any]:
        return {"phi": self.phi, "N": self.N, "Ndim": self.Ndim, "box": deepcopy(self.box),
                "walls": deepcopy(self.walls), "fix_height": self.fix_height,
                "dist": deepcopy(self.dist), "seed": self.seed}

    def _collect_pack_input(self) -> dict[str, Any]:
        return {"positions": deepcopy(self.positions), "diameters": deepcopy(self.diameters)}

    def _collect_pack_config(self) -> dict[str, Any]:
        return {"box": deepcopy(self.box), "walls": deepcopy(self.walls),
                "neighbor_max": self.neighbor_max, "seed": self.seed,
                "fix_height": self.fix_height}

    # ------------------------------------------------------------- actions
    def initialize(self) -> dict[str, Any]:
        self._log("rcpgenerator: initializing particles")
        result = _native.initialize_particles(self._collect_init_config())
        with self._internal():
            self.positions = result["positions"]
            self.diameters = result["diameters"]
            self.box = result["box"]
            self.walls = result["walls"]
            self._clear_results()
        self._set_flags(initialized=True, needs_initialize=False, needs_pack=True, packed=False)
        self._log("rcpgenerator: initialization complete")
        return result

    def _apply_result(self, result: dict[str, Any]) -> None:
        cf = _container_fraction(result["walls"])
        with self._internal():
            self.positions = result["positions"]
            self.diameters = result["diameters"]
            self.box = result["box"]
            self.walls = result["walls"]
            self.steps = result["steps"]
            self.phi_final = result["phi"] / cf
            self.max_min_dist = result["max_min_dist"]
            self.force_magnitude = result["force_magnitude"]
            hist = result["phi_history"]
            self.phi_history = [v / cf for v in hist] if cf != 1.0 else hist
            self.force_history = result["force_history"]
            self.energy_history = result["energy_history"]
            for key in ("max_overlap_history", "mu_flag_history", "mu_history", "alpha_history"):
                setattr(self, key, result.get(key, []))
        self._set_flags(packed=True, needs_pack=False)

    def _apply_trace(self, trace: dict[str, Any]) -> None:
        with self._internal():
            for key in _TRAJ_FIELDS:
                setattr(self, "trajectory_" + key, trace.get(key, []))

    def _execute_packing(self, *, verbose: bool, progress_interval: int,
                         capture_trajectory: bool, trajectory_interval: int,
                         options: dict[str, Any] | None = None) -> dict[str, Any]:
        self._clear_trajectory()
        if verbose:
            print("rcpgenerator: running packing")
        if verbose or capture_trajectory or options:
            callback: Callable | None = None
            if verbose:
                def callback(update: dict[str, Any]) -> None:
                    print(f"step={update['step']} phi={update['phi']:.6f} "
                          f"force={update['force_magnitude']:.6g} "
                          f"energy={update['energy']:.6g}")
            out = _native.run_packing_observed(
                self._collect_pack_input(), self._collect_pack_config(),
                progress_interval if verbose else 0, capture_trajectory,
                trajectory_interval if capture_trajectory else 0, callback, options or {})
            result = out["result"]
            self._apply_result(result)
            self._apply_trace(out["trace"])
        else:
            result = _native.run_packing(self._collect_pack_input(), self._collect_pack_config())
            self._apply_result(result)
        if verbose:
            print("rcpgenerator: packing complete")
        return result

    def _ensure_initialized(self, verbose: bool, why: str, rerun: str) -> None:
        if self._needs_initialize or not self._initialized:
            if verbose:
                print(f"rcpgenerator: initialization required before {why}")
            self.initialize()
        elif self._needs_pack and verbose:
            print(f"rcpgenerator: {rerun}")

    def pack(self, verbose: bool | None = None, progress_interval: int = 1000,
             capture_trajectory: bool = False, trajectory_interval: int = 1000) -> dict[str, Any]:
        v = self.verbose if verbose is None else verbose
        self._ensure_initialized(v, "packing", "rerunning packing on current realized state")
        return self._execute_packing(verbose=v, progress_interval=progress_interval,
                                     capture_trajectory=capture_trajectory,
                                     trajectory_interval=trajectory_interval)

    def relax(self, n_steps: int, mu: float | None = None, fix_diameter: bool = False,
              target_phi: float | None = None, verbose: bool = False,
              progress_interval: int = 1000, capture_trajectory: bool = False,
              trajectory_interval: int = 1000) -> dict[str, Any]:
        if n_steps <= 0:
            raise ValueError("n_steps must be positive")
        if target_phi is not None and fix_diameter:
            raise ValueError("target_phi cannot be combined with fix_diameter=True")
        self._ensure_initialized(verbose, "relaxation", "relaxing current realized state")
        options: dict[str, Any] = {"max_steps": int(n_steps), "fix_diameter": bool(fix_diameter)}
        if mu is not None:
            options["mu"] = float(mu)
        if target_phi is not None:
            options["target_phi"] = float(target_phi)
        return self._execute_packing(verbose=verbose, progress_interval=progress_interval,
                                     capture_trajectory=capture_trajectory,
                                     trajectory_interval=trajectory_interval, options=options)

    # ------------------------------------------------------ state updates
    def _current_phi(self) -> float:
        if not self.positions or not self.diameters or not self.box:
            raise ValueError("Current realized state is unavailable")
        box_volume = math.prod(self.box) if self.box else 1.0
        unit = math.pi ** (self.Ndim / 2.0) / math.gamma(self.Ndim / 2.0 + 1.0)
        volume = sum(unit * (d / 2.0) ** self.Ndim for d in self.diameters)
        modifier = _ball_fraction(-self.walls[0]) if self.walls and self.walls[0] < 0 else 1.0
        return volume / box_volume / modifier

    def _state_changed(self) -> None:
        self._clear_results()
        self._clear_trajectory()
        self._set_flags(needs_pack=True, packed=False)

    def update_phi(self, target_phi: float) -> float:
        if target_phi <= 0:
            raise ValueError("target_phi must be positive")
        exponent = self.Ndim - 1 if self.fix_height else self.Ndim
        scale = (target_phi / self._current_phi()) ** (1.0 / exponent)
        with self._internal():
            self.diameters = [d * scale for d in self.diameters]
            if self.fix_height:
                last = self.Ndim - 1
                box = list(self.box)
                box[last] *= scale
                self.box = box
                self.positions = [[c * scale if k == last else c for k, c in enumerate(p)]
                                  for p in self.positions]
        self._state_changed()
        return self._current_phi()

    def update_box(self, box: list[float]) -> list[float]:
        if len(box) != self.Ndim:
            raise ValueError("box length must match Ndim")
        if not self.positions or not self.diameters or not self.box:
            raise ValueError("Current realized state is unavailable")
        old = list(self.box)
        new = [float(v) for v in box]
        old_volume, new_volume = math.prod(old), math.prod(new)
        if old_volume <= 0 or new_volume <= 0:
            raise ValueError("box dimensions must be positive")
        d_scale = (new_volume / old_volume) ** (1.0 / self.Ndim)
        p_scale = [new[k] / old[k] for k in range(self.Ndim)]
        with self._internal():
            self.box = new
            self.positions = [[c * p_scale[k] for k, c in enumerate(p)] for p in self.positions]
            self.diameters = [d * d_scale for d in self.diameters]
        self._state_changed()
        return list(self.box)

    def update_height(self, height: float) -> list[float]:
        box = list(self.box)
        if not box:
            raise ValueError("Current box is unavailable")
        box[self.Ndim - 1] = float(height)
        return self.update_box(box)

    def reset(self) -> None:
        with self._internal():
            self.positions = []
            self.diameters = []
            self._clear_results()
            self._clear_trajectory()
        self._set_flags(initialized=False, packed=False, needs_initialize=True, needs_pack=True)

    def copy(self) -> "Packing":
        return type(self).from_dict(self.to_dict())

    def clone(self) -> "Packing":
        return self.copy()

    def summary(self) -> str:
        return (f"Packing summary: N={self.N}, Ndim={self.Ndim}, phi={self.phi}, "
                f"initialized={self.initialized}, packed={self.packed}, "
                f"needs_initialize={self.needs_initialize}, needs_pack={self.needs_pack}, "
                f"phi_final={self.phi_final}, steps={self.steps}, "
                f"force_magnitude={self.force_magnitude}, "
                f"trajectory_samples={len(self.trajectory_steps)}")

    def to_dict(self) -> dict[str, Any]:
        return {
            "init_config": self._collect_init_config(),
            "pack_config": self._collect_pack_config(),
            "realized_state": self._collect_pack_input(),
            "final_result": {
                "steps": self.steps, "phi_final": self.phi_final,
                "max_min_dist": self.max_min_dist, "force_magnitude": self.force_magnitude,
                "phi_history": deepcopy(self.phi_history),
                "force_history": deepcopy(self.force_history),
                "energy_history": deepcopy(self.energy_history),
            },
            "trajectory": {key: deepcopy(getattr(self, "trajectory_" + key))
                           for key in _TRAJ_FIELDS},
            "status": {"initialized": self._initialized, "packed": self._packed,
                       "needs_initialize": self._needs_initialize,
                       "needs_pack": self._needs_pack},
        }

    # ------------------------------------------------- visualisation (out)
    def render(self, path: str | Path | None = None, show: bool = False, palette_choice: int = 1,
               figsize=(4, 4), dpi=110):
        raise NotImplementedError("rendering is not part of the GPU engine; export positions "
                                  "with to_dict() and plot them with any tool")

    def savefig(self, path: str | Path, show: bool = False, palette_choice: int = 1):
        return self.render(path=path, show=show, palette_choice=palette_choice)

    def show_packing(self, palette_choice: int = 1, figsize=(4, 4), dpi=110) -> None:
        self.render(path=None, show=True, palette_choice=palette_choice, figsize=figsize, dpi=dpi)

    def animate_2d(self, path: str | Path | None = None, show: bool = True,
                   palette_choice: int = 1, interval_ms: int = 80, repeat: bool = False):
        raise NotImplementedError("animation is not part of the GPU engine")
