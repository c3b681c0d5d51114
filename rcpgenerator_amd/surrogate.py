"""phi_J / schedule surrogate for warm-started batched runs.

A small network (``capi.Mlp``: 3 hidden layers of 128, FP32, trained on the
GPU by this library's own forward and backward kernels) maps the diameter
distribution's moments, the size ratio, the dimension and the boundary to
two numbers: the jamming fraction phi_J the plain method reaches, and the
fraction of the default rung-1 deadline a warm-started run needs.  A warm
start (``warm_options``) begins the run just below the predicted phi_J
(``initial_phi_target``) with the shortened rung-1 deadline
(``rung1_deadline``, the reference's RCP_RUNG1_DEADLINE hook,
rcp_generator.cpp:5864-5868).  The plain growth and minimisation still
finish every packing; ``run_warm_batch`` judges each finished packing
against the prediction and reruns it with the default schedule when the
prediction missed.

Training rows are census rows (``rcpgenerator_amd.census``) with the
measured phi_J and the deadline-fraction label.
"""
from __future__ import annotations

import math
from typing import Sequence

import numpy as np

from . import capi
from .census import distribution_stats

N_FEATURES = 12
N_TARGETS = 2
HIDDEN = 128
LAYERS = 3


def default_deadline(N: int, ndim: int) -> float:
    """Rung-1 deadline of the "quick" preset (rcp_generator.cpp:5870-5885)."""
    return max(1500.0, 2.0 * 4000.0 + 9.0 * math.sqrt(float(N) * float(ndim)) - 2000.0)


def features(stats: dict, ndim: int, walls: Sequence[int], N: int) -> np.ndarray:
    """Network inputs of one packing: moments m2..m4 of D / mean(D) (as
    excesses over 1), log size ratio, one-hot dimension 2..5, the fraction
    of hard-wall axes, a curved-container flag and log10(N) / 5."""
    nd = int(ndim)
    w = [int(v) for v in walls] if len(walls) else [0] * nd
    curved = 1.0 if w and w[0] < 0 else 0.0
    hard = 0.0 if curved else float(sum(1 for v in w if v == 1)) / max(1, nd)
    f = np.zeros(N_FEATURES, dtype=np.float32)
    f[0] = stats["m2"] - 1.0
    f[1] = stats["m3"] - 1.0
    f[2] = stats["m4"] - 1.0
    f[3] = math.log(stats["d_max"] / stats["d_min"])
    if 2 <= nd <= 5:
        f[4 + nd - 2] = 1.0
    f[8] = hard
    f[9] = curved
    f[10] = math.log10(max(1, int(N))) / 5.0
    f[11] = float(nd) / 5.0
    return f


def features_of_packing(diameters, ndim, walls) -> np.ndarray:
    return features(distribution_stats(diameters), ndim, walls, len(diameters))


def features_of_row(row: dict) -> np.ndarray:
    return features(row, row["Ndim"], row.get("walls", []), row["N"])


class Surrogate:
    """The network plus its input / output standardisation."""

    def __init__(self, seed: int = 0):
        self.net = capi.Mlp(N_FEATURES, HIDDEN, LAYERS, N_TARGETS, seed=seed)
        self.x_mean = np.zeros(N_FEATURES, dtype=np.float32)
        self.x_std = np.ones(N_FEATURES, dtype=np.float32)
        self.y_mean = np.zeros(N_TARGETS, dtype=np.float32)
        self.y_std = np.ones(N_TARGETS, dtype=np.float32)

    def fit(self, X, Y, epochs: int = 300, batch: int = 128, lr: float = 1e-3,
            seed: int = 0) -> np.ndarray:
        X = np.asarray(X, dtype=np.float32)
        Y = np.asarray(Y, dtype=np.float32)
        self.x_mean = X.mean(axis=0)
        self.x_std = np.where(X.std(axis=0) > 1e-6, X.std(axis=0), 1.0).astype(np.float32)
        self.y_mean = Y.mean(axis=0)
        self.y_std = np.where(Y.std(axis=0) > 1e-6, Y.std(axis=0), 1.0).astype(np.float32)
        Xn = (X - self.x_mean) / self.x_std
        Yn = (Y - self.y_mean) / self.y_std
        return self.net.train(Xn, Yn, batch=batch, epochs=epochs, lr=lr, seed=seed)

    def predict(self, X) -> np.ndarray:
        X = np.asarray(X, dtype=np.float32).reshape(-1, N_FEATURES)
        Yn = self.net.predict((X - self.x_mean) / self.x_std)
        return Yn * self.y_std + self.y_mean

    def save(self, path) -> None:
        np.savez(path, params=self.net.get_params(), x_mean=self.x_mean, x_std=self.x_std,
                 y_mean=self.y_mean, y_std=self.y_std)

    @classmethod
    def load(cls, path) -> "Surrogate":
        d = np.load(path, allow_pickle=False)
        s = cls()
        s.net.set_params(d["params"])
        s.x_mean, s.x_std = d["x_mean"], d["x_std"]
        s.y_mean, s.y_std = d["y_mean"], d["y_std"]
        return s
