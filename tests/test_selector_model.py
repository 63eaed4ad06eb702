"""The shipped selector model (perf-libs-sparse_amd/selector/selector_mlp.txt,
trained on MI355X timings by the library's own MLP kernels) scored on the
held-out matrices measured on MI355X, with the numpy MLP of the oracle: regret
against the exhaustive autotuner within 10%, better than the hand heuristic
(the selector without the network) and faster than the fixed CSR-vector kernel."""
from __future__ import annotations

from pathlib import Path

import numpy as np

from oracle import mlp

ROOT = Path(__file__).resolve().parent.parent
SEL = ROOT / "perf-libs-sparse_amd" / "selector"


def load():
    from plsparse import selector as S

    params, mean, std = S.load_model(SEL / "selector_mlp.txt")
    d = np.load(SEL / "heldout_mi355x.npz")
    return S, params, mean, std, d


def test_model_file_shape():
    from plsparse import NUM_FEATURES, NUM_VARIANTS

    S, params, mean, std, _ = load()
    assert len(params) == mlp.num_params(NUM_VARIANTS)
    assert len(mean) == NUM_FEATURES and len(std) == NUM_FEATURES and np.all(std > 0)
    header = (SEL / "selector_mlp.txt").read_text().splitlines()[:2]
    assert header == ["plsparse-selector-mlp 1", f"features 12 hidden 64 outputs {NUM_VARIANTS}"]


def test_heldout_regret_and_ablation():
    S, params, mean, std, d = load()
    times = d["times"]
    K = times.shape[1]
    _, _, pred = mlp.forward(params, (d["features"] - mean) / std, K)
    learned = S.score(S.picks_of(pred, times), times)
    heuristic = S.score(d["heuristic"], times)
    fixed = S.score(np.zeros(len(times), np.int64), times)
    assert learned["n"] >= 100
    assert learned["regret"] <= 0.10
    assert learned["regret"] < heuristic["regret"]
    assert learned["speedup_vs_csr_vector"] > 1.0 and fixed["speedup_vs_csr_vector"] == 1.0
    # and below every fixed variant, the best of which (CSR-adaptive) is the
    # baseline a selector has to beat
    with np.errstate(divide="ignore"):  # variants that did not run on some matrix score inf
        for v in range(K):
            assert learned["regret"] < S.score(np.full(len(times), v), times)["regret"], v
    # held-out matrices come from other seeds: apart from stencils, whose
    # structure has no randomness, their features do not occur in training
    tr = np.load(SEL / "train_mi355x.npz")
    seen = {tuple(f) for f in tr["features"]}
    fresh = [tuple(f) not in seen for f, k in zip(d["features"], d["kinds"]) if k != "stencil"]
    assert all(fresh)


def test_score_definitions():
    from plsparse import selector as S

    times = np.array([[2.0, 1.0, -1.0], [1.0, 4.0, 2.0]])
    s = S.score(np.array([0, 0]), times)  # ratios 2 and 1
    assert abs(s["regret"] - (np.sqrt(2.0) - 1.0)) < 1e-12
    assert s["within_10pct"] == 0.5
    assert np.array_equal(S.picks_of(np.array([[0.3, 0.1, -5.0], [0.0, 0.2, 0.1]]), times), [1, 0])
    y, m = S.targets(times)
    assert np.allclose(y[0], [np.log(2.0), 0.0, 0.0]) and np.array_equal(m[0], [1, 1, 0])
