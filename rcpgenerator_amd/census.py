"""Census rows for packings run through this engine.

The rows follow the schema of rcptools' packing census
(rcptools/census.py:54-126): one JSON object per packing with the raw
moments m1..m6 of the realised, mean-normalised diameter distribution, its
extremes, and the run's outputs (phi, max overlap, steps, wall time).  The
same rows are the training records of the phi_J surrogate
(``rcpgenerator_amd.surrogate``).  Rows are appended as JSON lines, so an
interrupted campaign keeps what it wrote.
"""
from __future__ import annotations

import json
import os
from typing import Iterable, Sequence

import numpy as np

CENSUS_FILENAME = "packing_census.jsonl"
N_MOMENTS = 6


def distribution_stats(diameters) -> dict:
    """m1..m6 of D / mean(D), with the normalised d_min and d_max
    (rcptools/census.py:54-65 records the same statistics)."""
    D = np.asarray(diameters, dtype=np.float64)
    if D.size == 0 or not np.all(np.isfinite(D)) or D.mean() <= 0:
        raise ValueError("diameters must be finite with a positive mean")
    Dn = D / D.mean()
    out = {f"m{k}": float(np.mean(Dn ** k)) for k in range(1, N_MOMENTS + 1)}
    out["d_min"] = float(Dn.min())
    out["d_max"] = float(Dn.max())
    return out


def census_row(*, diameters, ndim: int, walls: Sequence[int], result: dict,
               search_name: str = "batch", model_name: str = "unnamed", candidate: int = 0,
               seed: int = 0, generation: int = 0, wall_s: float | None = None,
               theta: Iterable[float] = ()) -> dict:
    """One census row for a finished packing.  `result` needs phi,
    max_min_dist and steps (a run_packing dict or a batch summary);
    `diameters` are the packing's input diameters (their moments are
    scale-free).  A census row (max_overlap instead of max_min_dist) serves
    as `result` too."""
    stats = distribution_stats(diameters)
    phi = float(result["phi"])
    return {
        "search_name": search_name,
        "model_name": model_name,
        "generation": int(generation),
        "candidate": int(candidate),
        "seed": int(seed),
        "seed_ordinal": 0,
        "state": "done",
        "rejected": False,
        "N": int(len(np.asarray(diameters))),
        "Ndim": int(ndim),
        "size_ratio": stats["d_max"] / stats["d_min"],
        "phi": phi,
        "phi_corr": phi,
        "max_overlap": float(result["max_min_dist"] if "max_min_dist" in result
                             else result["max_overlap"]),
        "steps": int(result["steps"]) if result.get("steps") is not None else None,
        "wall_s": float(wall_s) if wall_s is not None else None,
        **stats,
        "theta": [float(t) for t in theta],
        "walls": [int(w) for w in walls],
    }


def census_path(save_dir) -> str:
    return os.path.join(str(save_dir), CENSUS_FILENAME)


def append_rows(path, rows: Iterable[dict]) -> int:
    """Append rows as JSON lines; returns how many were written."""
    n = 0
    with open(path, "a") as f:
        for r in rows:
            f.write(json.dumps(r) + "\n")
            n += 1
    return n


def load_rows(path) -> list:
    """Rows of a census file (blank lines skipped)."""
    out = []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line:
                out.append(json.loads(line))
    return out
