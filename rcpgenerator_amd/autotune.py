"""Tile autotuner for the tiled step kernel.

Times ``k_step_tile`` (HIP events around every step launch) and the
re-binning for a sweep of tile shapes -- the cell split (cells of 1/split of
the contact reach, stencils of +-split cells) and the target number of
particles per workgroup tile -- on one 2-D and one 3-D packing of 10^6
particles, steps 1..K, and writes the fastest shape per dimension as the
compiled-in table ``csrc/rcp_tile_table.h`` (the timings go into the
header's comment).  A shape is priced by step kernel plus re-binning time
per step: finer cells shorten the stencil scans but add cells to every
re-binning.  Run on the GPU:

    python -m rcpgenerator_amd.autotune --out rcpgenerator_amd/csrc/rcp_tile_table.h

then rebuild.  The engine never tunes at run time, so a run's tiling (and
the rounding of its fast-mode scalars) does not depend on timing noise.
"""
from __future__ import annotations

import argparse
import json
import math
import time

CANDIDATES = (64, 96, 128, 192, 256, 384)
SPLITS = (1, 2)
DEFAULT = (1, 128)   # (split, particles per tile), kept unless another shape is MARGIN faster
MARGIN = 0.03

WORKLOADS = {
    2: {"phi": 0.10, "N": 1_000_000, "Ndim": 2, "box": [1.0, 1.0], "walls": [0, 0], "seed": 123,
        "dist": {"type": "lognormal", "mu": -0.5 * math.log(1.04),
                 "sigma": math.sqrt(math.log(1.04))}},
    3: {"phi": 0.10, "N": 1_000_000, "Ndim": 3, "box": [1.0, 1.0, 1.0], "walls": [0, 0, 0],
        "seed": 123, "dist": {"type": "bidisperse", "d1": 1.0, "d2": 1.4, "p": 0.5}},
}


def sweep(nd: int, steps: int, log=print) -> dict:
    """{(split, per_tile): timings} for the workload of dimension nd."""
    from . import capi
    x, D, box, walls = capi.initialize(WORKLOADS[nd])
    out = {}
    for sp in SPLITS:
        for t in CANDIDATES:
            e = capi.Engine(x, D, box=box, walls=walls, max_steps=10 ** 6, time_kernels=True,
                            tile_particles=t, tile_split=sp)
            e.sync()
            t0 = time.perf_counter()
            e.run(steps)
            e.sync()
            wall = time.perf_counter() - t0
            tm = e.timing()
            st = e.stats()
            r = {"kernel_ms": tm["step_ms"] / max(tm["step_launches"], 1),
                 "wall_ms_per_step": 1e3 * wall / steps,
                 "rebin_ms_per_step": tm["build_ms"] / steps, "tile_steps": st["tile_steps"]}
            r["cost_ms"] = r["kernel_ms"] + r["rebin_ms_per_step"]
            out[(sp, t)] = r
            log(f"[autotune] ndim {nd} split {sp} tile {t}: kernel {r['kernel_ms']:.4f} ms, "
                f"re-binning {r['rebin_ms_per_step']:.4f} ms/step, "
                f"step {r['wall_ms_per_step']:.4f} ms")
            del e
    return out


HEADER = """// Tile shape of the tiled step per dimension -- cell split (cells of 1/split
// of the contact reach, stencils of +-split cells) and particles per
// workgroup tile -- as measured by the tile autotuner (python -m
// rcpgenerator_amd.autotune on one MI355X: 10^6 particles, steps 1..{steps};
// average k_step_tile launch plus re-binning per step, in ms, per
// split/particles):
{timings}// A shape replaces the default (split 1, 128 particles) only when it is at
// least 3 % faster.  The table is compiled in so that a run's tiling, and with
// it the rounding of its fast-mode scalars, does not depend on timing noise
// at run time.
#pragma once

namespace rcp {{

struct TileTuning {{
    int ndim;
    int per_tile;   // target particles per tile (tile_len = per_tile / particles per cell)
    int split;      // cells of 1/split of the contact reach
}};

constexpr TileTuning kTileTable[] = {{
{rows}}};

inline int tuned_tile_particles(int ndim)
{{
    for (const TileTuning& t : kTileTable)
        if (t.ndim == ndim) return t.per_tile;
    return 128;
}}

inline int tuned_tile_split(int ndim)
{{
    for (const TileTuning& t : kTileTable)
        if (t.ndim == ndim) return t.split;
    return 1;
}}

}}  // namespace rcp
"""


def choose(r: dict) -> tuple:
    """The cheapest (split, per_tile) shape, unless it beats DEFAULT by less
    than MARGIN (timing noise must not move the tiling)."""
    best = min(r, key=lambda k: r[k]["cost_ms"])
    if DEFAULT in r and r[best]["cost_ms"] > (1.0 - MARGIN) * r[DEFAULT]["cost_ms"]:
        return DEFAULT
    return best


def _key(k: str) -> tuple:
    sp, t = k.split("/")
    return int(sp), int(t)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m rcpgenerator_amd.autotune")
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--out", required=True, help="header to write")
    ap.add_argument("--results", help="JSON of an earlier sweep (its last line): write the "
                                      "header from it instead of measuring")
    a = ap.parse_args(argv)
    best, timings = {}, ""
    results = {}
    if a.results:
        with open(a.results) as f:
            last = [ln for ln in f if ln.startswith("{")][-1]
        results = {int(nd): {_key(k): v for k, v in r.items()}
                   for nd, r in json.loads(last)["results"].items()}
    for nd in (2, 3):
        r = results[nd] if a.results else sweep(nd, a.steps, log=lambda m: print(m, flush=True))
        results[nd] = r
        best[nd] = choose(r)
        for sp in sorted({k[0] for k in r}):
            timings += f"//   {nd}-D split {sp}: " + ", ".join(
                f"{t}: {r[(sp, t)]['cost_ms']:.4f}" for (s2, t) in sorted(r) if s2 == sp) + "\n"
    rows = "".join(f"    {{{nd}, {best[nd][1]}, {best[nd][0]}}},\n" for nd in (2, 3))
    with open(a.out, "w") as f:
        f.write(HEADER.format(steps=a.steps, timings=timings, rows=rows))
    print(json.dumps({"best": {nd: list(b) for nd, b in best.items()},
                      "results": {nd: {f"{k[0]}/{k[1]}": v for k, v in r.items()}
                                  for nd, r in results.items()}}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
