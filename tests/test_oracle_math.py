"""Pins the oracle's formulas with mathematics the reference does not state
as tests: closed forms for one contact of each kind, Newton's third law,
F = -grad E and dkappa = -dE/dkappa by finite differences, the first ADAM
step, and brute-force reductions.  A dropped term, a wrong sign, a swapped
index or a missing factor fails one of these."""
from __future__ import annotations

import math

import numpy as np
import pytest


def energy(x, D, box, walls):
    """E = sum_pairs (r - d)^2 / 2 + sum_flat (r - d)^2 + sum_curved (r - d')^2,
    minimum image on every axis, r the sum of radii (walls: the radius)."""
    N, nd = x.shape
    E = 0.0
    for i in range(N):
        for j in range(i + 1, N):
            dx = x[j] - x[i]
            dx -= np.rint(dx / box) * box
            d = math.sqrt(float(dx @ dx))
            r = 0.5 * (D[i] + D[j])
            if d < r:
                E += 0.5 * (r - d) ** 2
        for k in range(nd):
            if walls[k] == 1:
                r = D[i] / 2
                for c in (0.0, box[k]):
                    dist = abs(c - x[i, k])
                    if dist < r:
                        E += (r - dist) ** 2
        if walls[0] < 0:
            t = -walls[0]
            R = box[0] / 2
            rho = math.sqrt(float(np.sum((x[i, :t] - R) ** 2)))
            r = D[i] / 2
            if R - rho < r:
                E += (r - (R - rho)) ** 2
    return E


def test_single_pair_closed_form(oracle):
    x = np.array([[0.40, 0.50], [0.55, 0.50]])
    D = np.array([0.20, 0.20])
    F, s = oracle.forces(x, D, np.ones(2), np.zeros(2), 0.0)
    d, r = 0.15, 0.20
    f = -(r / d - 1.0)                      # F_i += f * (x_j - x_i)
    assert F[0, 0] == pytest.approx(f * 0.15, rel=1e-14)
    assert F[1, 0] == pytest.approx(-f * 0.15, rel=1e-14)
    assert F[0, 1] == 0.0 and F[1, 1] == 0.0
    ov = 1 - d / r
    assert s["count"] == 2                  # count starts at one
    assert s["U"] == pytest.approx((ov / 2) ** 2, rel=1e-14)
    assert s["max_min_dist"] == pytest.approx(ov, rel=1e-14)
    assert s["Lc"] == pytest.approx(r / 2, rel=1e-14)
    assert s["Fmean"] == pytest.approx(-f / 2, rel=1e-14)
    assert s["dkappa"] == pytest.approx(r * (d - r) + 0.0, rel=1e-14)


def test_mu_term_of_dkappa(oracle):
    x = np.array([[0.2, 0.2], [0.7, 0.7]])
    D = np.array([0.1, 0.3])
    _, s0 = oracle.forces(x, D, np.ones(2), np.zeros(2), 0.0)
    _, s1 = oracle.forces(x, D, np.ones(2), np.zeros(2), 2.0)
    assert s0["dkappa"] == 0.0
    assert s1["dkappa"] == pytest.approx(2.0 * 0.4, rel=1e-15)


def test_flat_wall_closed_form(oracle):
    x = np.array([[0.03, 0.5]])
    D = np.array([0.10])
    F, s = oracle.forces(x, D, np.ones(2), np.array([1, 0]), 0.0)
    r, dist = 0.05, 0.03
    f = -2.0 * (r / dist - 1.0)
    assert F[0, 0] == pytest.approx(f * (0.0 - 0.03), rel=1e-14)   # pushed to +x
    assert F[0, 0] > 0
    assert s["U"] == pytest.approx((2.0 * (1 - dist / r) / 2) ** 2, rel=1e-14)
    assert s["dkappa"] == pytest.approx(2 * r * (dist - r), rel=1e-14)


def test_curved_container_closed_form(oracle):
    # disk of diameter 1 centred at 0.5; particle protruding by 0.02
    x = np.array([[0.5 + 0.47, 0.5]])
    D = np.array([0.10])
    F, s = oracle.forces(x, D, np.ones(2), np.array([-2, -1]), 0.0)
    r, gap = 0.05, 0.5 - 0.47
    f = -2.0 * abs(r / gap - 1.0)
    assert F[0, 0] == pytest.approx(f * gap, rel=1e-12)            # pushed inward
    assert F[0, 0] < 0 and F[0, 1] == 0.0
    assert s["U"] == pytest.approx(((1 - gap / r) / 2) ** 2, rel=1e-12)  # no factor 2


def _random(rng, N, nd, phi):
    D = np.exp(rng.normal(0, 0.2, N))
    unit = math.pi ** (nd / 2) / math.gamma(nd / 2 + 1)
    D *= (phi / (unit * np.sum((D / 2) ** nd))) ** (1 / nd)
    return rng.random((N, nd)), D


@pytest.mark.parametrize("nd", [2, 3, 4])
def test_newton_third_law(oracle, nd):
    rng = np.random.default_rng(nd)
    x, D = _random(rng, 200, nd, 0.7 if nd == 2 else 0.5)
    F, s = oracle.forces(x, D, np.ones(nd), np.zeros(nd), 0.0)
    assert s["count"] > 50
    scale = np.abs(F).sum()
    assert np.abs(F.sum(axis=0)).max() <= 1e-13 * scale


@pytest.mark.parametrize("walls", [[0, 0], [1, 1], [-2, -1], [0, 0, 1], [-3, -1, -1]])
def test_force_is_minus_energy_gradient(oracle, walls):
    nd = len(walls)
    rng = np.random.default_rng(7 + nd)
    x, D = _random(rng, 40, nd, 0.6 if nd == 2 else 0.4)
    if walls[0] < 0:
        # keep every centre inside the ball (the energy form assumes rho < R)
        t = -walls[0]
        for i in range(len(x)):
            while np.sum((x[i, :t] - 0.5) ** 2) > 0.46 ** 2:
                x[i, :t] = rng.random(t)
    box = np.ones(nd)
    w = np.array(walls, dtype=np.int8)
    F, s = oracle.forces(x, D, box, w, 0.0)
    h = 1e-7
    for i in range(0, 40, 5):
        for k in range(nd):
            xp, xm = x.copy(), x.copy()
            xp[i, k] += h
            xm[i, k] -= h
            g = (energy(xp, D, box, w) - energy(xm, D, box, w)) / (2 * h)
            assert F[i, k] == pytest.approx(-g, abs=2e-7, rel=1e-5)
    # dkappa (without mu) = -dE/dkappa at kappa = 1 for D = D0 * kappa
    e = 1e-7
    g = (energy(x, D * (1 + e), box, w) - energy(x, D * (1 - e), box, w)) / (2 * e)
    assert s["dkappa"] == pytest.approx(-g, rel=1e-5, abs=1e-9)


def test_mean_force_brute_force(oracle):
    rng = np.random.default_rng(1)
    F = rng.normal(size=(500, 3))
    F[::5] = 0.0
    F[1::7] = 1e-9                           # |F|^2 = 3e-18 < 1e-16: excluded
    mags = np.sqrt((F ** 2).sum(axis=1))
    keep = (F ** 2).sum(axis=1) > 1e-16
    expect = mags[keep].mean() / math.sqrt(3)
    assert oracle.mean_force(F) == pytest.approx(expect, rel=1e-13)
    assert math.isnan(oracle.mean_force(np.zeros((4, 2))))


def test_adam_first_step_known_answer(oracle):
    rng = np.random.default_rng(3)
    F = rng.normal(size=(300, 2)) * 1e-3
    x = rng.random((300, 2)) * 0.5 + 0.25
    z = np.zeros_like(F)
    alpha = 0.005
    m, v, x1 = oracle.adam_positions(F, z, z, x, np.ones(2), alpha, 1 / (1 - 0.9), 1 / (1 - 0.999))
    # m1 = -(1-b1) F, v1 = (1-b2) F^2  ->  m_hat = -F, v_hat = F^2
    np.testing.assert_allclose(m, -0.1 * F, rtol=1e-14)
    np.testing.assert_allclose(v, 0.001 * F * F, rtol=1e-14)
    np.testing.assert_allclose(x1 - x, alpha * F / (np.abs(F) + 1e-8), rtol=1e-9, atol=1e-16)


def test_adam_wrap_into_box(oracle):
    F = np.array([[1.0, -1.0]])
    x = np.array([[0.999, 0.001]])
    _, _, x1 = oracle.adam_positions(F, np.zeros((1, 2)), np.zeros((1, 2)), x, np.ones(2), 0.005,
                                     10.0, 1000.0)
    assert 0.0 <= x1[0, 0] < 1.0 and 0.0 <= x1[0, 1] < 1.0
    assert x1[0, 0] == pytest.approx(0.999 + 0.005 / (1 + 1e-8) - 1.0, abs=1e-12)


@pytest.mark.parametrize("nd", [2, 3])
def test_reorder_permutation_sorts_keys(oracle, nd):
    rng = np.random.default_rng(11)
    x = rng.random((2000, nd))
    perm = oracle.reorder_perm(x, np.ones(nd))
    assert sorted(perm.tolist()) == list(range(2000))
    if nd == 3:
        q = np.minimum(np.maximum(x, 0.0), 1.0 - 1e-12) * float(2 ** 21 - 1)
        q = q.astype(np.uint64)
        code = np.zeros(2000, dtype=np.uint64)
        for b in range(21):
            for d in range(3):
                code |= ((q[:, d] >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b + d)
        keys = code[perm]
    else:
        keys = x[perm, 0]
    assert np.all(keys[1:] >= keys[:-1])


def test_scale_diameters_hits_target(oracle):
    rng = np.random.default_rng(2)
    for nd in (2, 3, 4, 5):
        D = rng.random(100) + 0.5
        box = np.array([1.0, 1.3, 0.8, 1.1, 0.9][:nd])
        out, f = oracle.scale_diameters(D, 0.4, box, nd)
        unit = math.pi ** (nd / 2) / math.gamma(nd / 2 + 1)
        assert unit * np.sum((out / 2) ** nd) / np.prod(box) == pytest.approx(0.4, rel=1e-12)
        np.testing.assert_allclose(out / D, f, rtol=1e-15)


def test_pair_set_brute_force_predicate(oracle):
    """Full-refresh set = every i<j with d < (Di+Dj)/2 + t_shell (kd-tree test)."""
    rng = np.random.default_rng(5)
    for nd in (2, 3):
        x, D = _random(rng, 300, nd, 0.4)
        off, nbr = oracle.pairs_full(x, D, np.ones(nd))
        t = oracle.shell(D)
        s = np.sort(D)
        assert t == pytest.approx(min(max(s[150] * 1.02, s[0] * 2.25), s[-1]) * 0.5)
        got = set()
        for i in range(300):
            for e in range(int(off[i]), int(off[i + 1])):
                got.add((i, int(nbr[e])))
        want = set()
        for i in range(300):
            dx = x[i + 1:] - x[i]
            dx -= np.rint(dx)
            d = np.sqrt((dx ** 2).sum(axis=1))
            rc = (D[i] + D[i + 1:]) * 0.5 + t
            for j in np.nonzero(d < rc)[0]:
                want.add((i, i + 1 + int(j)))
        assert got == want
