"""The phi_J surrogate trained on the GPU (rcpgenerator_amd.surrogate) from
the committed training rows (rcpgenerator_amd/data/surrogate_train.npz:
plain and warm-started batched runs of 640 config-4-style packings, N
500..2000, 2-4 dimensions), and warm-started batched runs guided by it."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DATA = os.path.join(os.path.dirname(__file__), "..", "rcpgenerator_amd", "data")


def test_network_beats_the_dimension_mean():
    """What the trained network gains: on rows it never saw, its phi_J error
    is well below that of the per-dimension mean, the predictor without it."""
    from rcpgenerator_amd import surrogate as S
    d = np.load(os.path.join(DATA, "surrogate_train.npz"), allow_pickle=False)
    X, Y, nd = d["X"], d["Y"], d["meta"][:, 2]
    order = np.random.default_rng(7).permutation(len(X))
    te, tr = order[:128], order[128:]
    s = S.Surrogate(seed=3)
    s.fit(X[tr], Y[tr], epochs=300)
    P = s.predict(X[te])
    mae_net = np.mean(np.abs(P[:, 0] - Y[te, 0]))
    means = {k: Y[tr][nd[tr] == k, 0].mean() for k in (2, 3, 4)}
    mae_base = np.mean(np.abs(np.array([means[k] for k in nd[te]]) - Y[te, 0]))
    assert mae_net < 0.6 * mae_base, (mae_net, mae_base)
    # the safe-schedule label: at least as good as always answering the majority
    acc = np.mean((P[:, 1] >= 0.5) == (Y[te, 1] >= 0.5))
    assert acc >= max(np.mean(Y[te, 1] >= 0.5), np.mean(Y[te, 1] < 0.5))


def test_save_load_round_trip(tmp_path):
    from rcpgenerator_amd import surrogate as S
    d = np.load(os.path.join(DATA, "surrogate_train.npz"), allow_pickle=False)
    s = S.Surrogate(seed=0)
    s.fit(d["X"][:200], d["Y"][:200], epochs=10)
    p = str(tmp_path / "m.npz")
    s.save(p)
    t = S.Surrogate.load(p)
    assert np.array_equal(s.predict(d["X"][200:260]), t.predict(d["X"][200:260]))


@pytest.mark.skipif(not os.path.exists(os.path.join(DATA, "surrogate.npz")),
                    reason="no trained model committed")
def test_guided_batch_keeps_phi_and_saves_steps_where_warm():
    """Guided batched runs of held-out 2-D packings: every packing jams,
    phi stays within 0.005 of the plain method, and the packings the network
    warm-started (and did not rerun) needed fewer steps."""
    from rcpgenerator_amd import surrogate as S
    from rcpgenerator_amd.workloads import batch_cases
    surr = S.Surrogate.load(os.path.join(DATA, "surrogate.npz"))
    cases = batch_cases(8, 700000, 600, 1200, (2,))
    out = S.evaluate(surr, cases, log=lambda m: None, detail=True)
    assert out["jammed_guided"] == 1.0
    assert out["phi_max_abs_diff"] <= 0.005
    warm = [k for k in out["warm_ids"] if k not in out["rerun_ids"]]
    assert warm, "the network warm-started none of the 2-D packings"
    for k in warm:
        assert out["steps_guided_each"][k] < out["steps_plain_each"][k]
