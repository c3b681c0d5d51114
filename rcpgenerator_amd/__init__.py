"""MI355X (gfx950, HIP) engine for random-close-packing generation.

Drop-in for the hot path behind ``rcpgenerator.Packing.pack()``
(KD-physics/RCPGenerator): the same ``Packing`` API and result dictionary,
and a ``_rcpgenerator`` module with the reference binding's five functions,
backed by hand-written HIP kernels.  There is no CPU fallback: importing the
package fails loudly when the native extension has not been built.

    import rcpgenerator_amd as rcpgenerator
    p = rcpgenerator.Packing(phi=0.08, N=2000, Ndim=3, box=[1, 1, 1])
    p.pack()
"""
from __future__ import annotations

import os as _os

_HERE = _os.path.dirname(_os.path.abspath(__file__))

try:
    from . import _rcpgenerator  # noqa: F401  (native module, built in-tree)
except ImportError as _e:  # pragma: no cover - exercised only on a broken tree
    raise ImportError(
        "rcpgenerator_amd: the native extension is not built. Run "
        "`python -c 'import __graft_entry__ as g; g.build()'` or "
        "`python rcpgenerator_amd/_build.py` in the repository root."
    ) from _e

from ._rcpgenerator import (  # noqa: E402
    Engine,
    get_num_threads,
    initialize_particles,
    last_batch_stats,
    run_packing,
    run_packing_batch,
    run_packing_observed,
    set_num_threads,
)
from ._wrapper import Packing  # noqa: E402

__all__ = [
    "Engine",
    "Packing",
    "get_num_threads",
    "initialize_particles",
    "last_batch_stats",
    "run_packing",
    "run_packing_batch",
    "run_packing_observed",
    "set_num_threads",
]


def native_library_path() -> str:
    """Path of the loaded HIP library (for load checks)."""
    return _os.path.join(_HERE, "lib", "librcp_amd.so")
