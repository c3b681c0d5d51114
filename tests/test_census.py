"""Census rows (rcpgenerator_amd.census): the reference census schema
(rcptools/census.py:96-121), scale-free moments and JSON-lines round trips.
CPU only."""
import numpy as np
import pytest

from rcpgenerator_amd import census

# keys of one row of the reference census (rcptools/census.py:96-121)
REFERENCE_KEYS = {"search_name", "model_name", "generation", "candidate", "seed",
                  "seed_ordinal", "state", "rejected", "N", "Ndim", "size_ratio", "phi",
                  "phi_corr", "max_overlap", "steps", "wall_s", "m1", "m2", "m3", "m4", "m5",
                  "m6", "d_min", "d_max", "theta"}


def test_moments_are_scale_free_and_normalised():
    r = np.random.default_rng(0)
    D = r.lognormal(0.0, 0.3, 1000)
    a, b = census.distribution_stats(D), census.distribution_stats(7.5 * D)
    for k in a:
        assert a[k] == pytest.approx(b[k], rel=1e-12)
    assert a["m1"] == pytest.approx(1.0, rel=1e-14)
    # Jensen: E[X^k] grows with k for a non-degenerate X with E[X] = 1
    assert 1.0 < a["m2"] < a["m3"] < a["m4"] < a["m5"] < a["m6"]
    assert a["d_min"] == pytest.approx(D.min() / D.mean()) and a["d_max"] == pytest.approx(D.max() / D.mean())


def test_mono_and_two_point_distributions():
    m = census.distribution_stats(np.full(10, 2.0))
    assert all(m[f"m{k}"] == 1.0 for k in range(1, 7)) and m["d_min"] == m["d_max"] == 1.0
    # half 1.0, half 3.0: D / mean = 0.5 and 1.5
    t = census.distribution_stats(np.array([1.0, 3.0] * 4))
    assert t["m2"] == pytest.approx((0.25 + 2.25) / 2) and t["m3"] == pytest.approx((0.125 + 3.375) / 2)


def test_row_schema_and_jsonl_round_trip(tmp_path):
    row = census.census_row(diameters=[1.0, 1.4, 1.0, 1.4], ndim=3, walls=[0, 0, 1],
                            result={"phi": 0.64, "max_min_dist": 1e-4, "steps": 12000},
                            search_name="s", model_name="bidisperse", candidate=3, seed=9,
                            wall_s=1.5)
    assert REFERENCE_KEYS <= set(row)
    assert row["size_ratio"] == pytest.approx(1.4) and row["N"] == 4 and row["steps"] == 12000
    p = census.census_path(tmp_path)
    assert census.append_rows(p, [row, row]) == 2
    assert census.append_rows(p, [row]) == 1
    back = census.load_rows(p)
    assert len(back) == 3 and back[0] == row


def test_bad_diameters_raise():
    with pytest.raises(ValueError):
        census.distribution_stats([])
    with pytest.raises(ValueError):
        census.distribution_stats([1.0, float("nan")])
