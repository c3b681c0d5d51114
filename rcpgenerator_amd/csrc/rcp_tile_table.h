// Particles per workgroup tile of the tiled step, per dimension, as measured
// by the tile autotuner (python -m rcpgenerator_amd.autotune on one MI355X:
// 10^6 particles, steps 1..400; average k_step_tile launch in ms):
//   2-D: 64: 0.2718, 96: 0.2125, 128: 0.2156, 192: 0.2766, 256: 0.3381, 384: 0.4657
//   3-D: 64: 0.8198, 96: 0.9519, 128: 1.1854, 192: 0.9115, 256: 0.9085, 384: 0.9088
// A size replaces the default 128 only when it is at least 3 % faster.
// The table is compiled in so that a run's tiling, and with it the rounding
// of its fast-mode scalars, does not depend on timing noise at run time.
#pragma once

namespace rcp {

struct TileTuning {
    int ndim;
    int per_tile;   // target particles per tile (tile_len = per_tile / particles per cell)
};

constexpr TileTuning kTileTable[] = {
    {2, 128},
    {3, 64},
};

inline int tuned_tile_particles(int ndim)
{
    for (const TileTuning& t : kTileTable)
        if (t.ndim == ndim) return t.per_tile;
    return 128;
}

}  // namespace rcp
