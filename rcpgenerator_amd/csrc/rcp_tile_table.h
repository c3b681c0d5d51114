// Tile shape of the tiled step per dimension -- cell split (cells of 1/split
// of the contact reach, stencils of +-split cells) and particles per
// workgroup tile -- as measured by the tile autotuner (python -m
// rcpgenerator_amd.autotune on one MI355X: 10^6 particles, steps 1..400;
// average k_step_tile launch plus re-binning per step, in ms, per
// split/particles):
//   2-D split 1: 64: 0.3837, 96: 0.3220, 128: 0.3264, 192: 0.3867, 256: 0.4497, 384: 0.5776
//   2-D split 2: 64: 0.3827, 96: 0.3992, 128: 0.4296, 192: 0.6708, 256: 0.6339, 384: 0.8920
//   3-D split 1: 64: 0.9532, 96: 1.0839, 128: 1.3306, 192: 1.0369, 256: 1.0354, 384: 1.0350
//   3-D split 2: 64: 4.1918, 96: 4.1934, 128: 4.1945, 192: 4.1925, 256: 4.1927, 384: 4.1955
// A shape replaces the default (split 1, 128 particles) only when it is at
// least 3 % faster.  The table is compiled in so that a run's tiling, and with
// it the rounding of its fast-mode scalars, does not depend on timing noise
// at run time.
#pragma once

namespace rcp {

struct TileTuning {
    int ndim;
    int per_tile;   // target particles per tile (tile_len = per_tile / particles per cell)
    int split;      // cells of 1/split of the contact reach
};

constexpr TileTuning kTileTable[] = {
    {2, 128, 1},
    {3, 64, 1},
};

inline int tuned_tile_particles(int ndim)
{
    for (const TileTuning& t : kTileTable)
        if (t.ndim == ndim) return t.per_tile;
    return 128;
}

inline int tuned_tile_split(int ndim)
{
    for (const TileTuning& t : kTileTable)
        if (t.ndim == ndim) return t.split;
    return 1;
}

}  // namespace rcp
