"""CPU checks for the learned FFT-length selector.

The float64 MLP oracle (oracle/mlp.py) is pinned by mathematics, not by
restating it: central finite differences of the loss, the closed-form
least-squares gradient of a single linear layer, tanh's derivative through a
one-hidden-unit network written out by hand, and Adam's first step, which is
lr * sign(g) up to eps.  Host-side selector pieces (features, candidate
lengths, model-file grammar, the round-off check) run through the C-ABI
library without a GPU.
"""
import math
import os

import numpy as np
import pytest

import cudalucas_amd as L
from cudalucas_amd import selector as S
from oracle import mlp as OM


def test_layout_counts_weights_and_biases():
    lay, total = OM.layout([7, 64, 64, 64, 1])
    assert total == 7 * 64 + 64 + 2 * (64 * 64 + 64) + 64 + 1
    assert lay[0] == (0, 448, 7, 64)


@pytest.mark.parametrize("sizes", [[3, 5, 1], [7, 16, 16, 1], [2, 1]])
def test_oracle_gradient_matches_finite_differences(sizes):
    rng = np.random.default_rng(len(sizes))
    _, total = OM.layout(sizes)
    params = rng.standard_normal(total) * 0.5
    x = rng.standard_normal((9, sizes[0]))
    t = rng.standard_normal(9)
    _, g = OM.loss_and_grad(params, sizes, x, t)
    h = 1e-6
    for i in range(total):
        e = np.zeros(total)
        e[i] = h
        lp, _ = OM.loss_and_grad(params + e, sizes, x, t)
        lm, _ = OM.loss_and_grad(params - e, sizes, x, t)
        assert g[i] == pytest.approx((lp - lm) / (2 * h), rel=1e-5, abs=1e-8), i


def test_oracle_linear_layer_is_least_squares():
    """One linear layer: L = mean((Xw + b - t)^2), dL/dw = 2/B X^T r."""
    rng = np.random.default_rng(11)
    X = rng.standard_normal((20, 4))
    t = rng.standard_normal(20)
    w = rng.standard_normal(4)
    b = 0.3
    r = X @ w + b - t
    loss, g = OM.loss_and_grad(np.concatenate([w, [b]]), [4, 1], X, t)
    assert loss == pytest.approx(np.mean(r ** 2))
    assert np.allclose(g[:4], 2.0 / 20 * X.T @ r)
    assert g[4] == pytest.approx(2.0 / 20 * r.sum())


def test_oracle_one_hidden_unit_by_hand():
    """y = v tanh(u x + a) + c; dL/du = 2/B sum r v (1 - tanh^2) x."""
    u, a, v, c = 0.7, -0.2, 1.3, 0.1
    x = np.array([[0.5], [-1.0], [2.0]])
    t = np.array([0.3, -0.4, 0.9])
    h = np.tanh(u * x[:, 0] + a)
    r = v * h + c - t
    loss, g = OM.loss_and_grad(np.array([u, a, v, c]), [1, 1, 1], x, t)
    d = 2.0 / 3 * r
    assert loss == pytest.approx(np.mean(r ** 2))
    assert g[0] == pytest.approx(np.sum(d * v * (1 - h ** 2) * x[:, 0]))
    assert g[1] == pytest.approx(np.sum(d * v * (1 - h ** 2)))
    assert g[2] == pytest.approx(np.sum(d * h))
    assert g[3] == pytest.approx(np.sum(d))


def test_oracle_adam_first_step_is_sign_of_gradient():
    g = np.array([3.0, -0.5, 1e-3])
    p, m, v = OM.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.01)
    assert np.allclose(p, -0.01 * np.sign(g), rtol=1e-3)  # eps shifts the smallest by 3e-4
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


# ------------------------------------------------------------ host side

def test_features_of_known_lengths():
    f = S.features(77005277, 4194304)
    assert f[0] == pytest.approx(math.log2(77005277), rel=1e-6)
    assert f[1] == 22 and f[2] == pytest.approx(77005277 / 4194304, rel=1e-6)
    assert list(f[3:]) == [0, 0, 0, 0]
    f = S.features(100003, 6144)        # 3 * 2^11, LDS-resident
    assert list(f[3:]) == [1, 0, 0, 1]
    f = S.features(2000003, 100 * 1024)  # 5^2 * 2^12, two sweeps
    assert list(f[3:]) == [0, 2, 0, 0]
    f = S.features(3000017, 3584 * 1024)  # 7 * 2^19
    assert list(f[3:]) == [0, 0, 1, 0]
    with pytest.raises(L.LLError):
        S.features(100003, 1000)


@pytest.mark.parametrize("p", [100_003, 523_007, 4_999_999, 20_000_003, 77_005_277, 119_999_987])
def test_candidates_bracket_the_reference_length(p):
    nref = L.choose_fft_length(p)[0]
    c = S.candidates(p)
    assert nref in c and c == sorted(c)
    assert all(L.fft_supported(n) and n % 1024 == 0 for n in c)
    assert all(0.62 * nref <= n <= 1.25 * nref and p / n <= 28 for n in c)


def test_careful_test_thresholds():
    """CUDALucas.cu:2616, 2629: any iteration > 0.35, or mean > 0.25, fails."""
    ok = np.full(1000, 0.2)
    assert S.careful_test(ok)
    bad = ok.copy()
    bad[500] = 0.3501
    assert not S.careful_test(bad)
    assert S.careful_test(np.full(1000, 0.25))
    assert not S.careful_test(np.full(1000, 0.2501))


def test_targets_clip_and_saturate():
    w = np.array([[0.1, 0.2], [0.7, 0.1], [np.nan, 0.0], [0.0, 0.0]])
    y = S.targets(w)
    assert y[0] == pytest.approx(math.log2(0.2))
    assert y[1] == y[2] == -1.0
    assert y[3] == -24.0


def test_growth_is_zero_at_the_window_and_increasing():
    assert S.growth(4096, 128, 128) == pytest.approx(0.0)
    g = S.growth(4096, np.array([1e3, 1e5, 1e7]), 128)
    assert (np.diff(g) > 0).all() and g[-1] < 0.5


def test_model_file_round_trip(tmp_path):
    _, total = OM.layout(S.SIZES)
    rng = np.random.default_rng(3)
    model = {"sizes": S.SIZES, "feature_mean": list(rng.standard_normal(7)),
             "feature_scale": list(rng.uniform(0.5, 2, 7)), "target": [-4.5, 2.25], "window": 128,
             "margin": 0.4, "limit": 0.35, "params": list(rng.standard_normal(total))}
    path = str(tmp_path / "m.txt")
    S.save_model(model, path, comment="test model")
    back = S.read_model(path)
    assert back["sizes"] == S.SIZES
    for k in ("feature_mean", "feature_scale", "target", "params"):
        assert np.allclose(back[k], model[k], rtol=1e-8)
    assert back["window"] == 128 and back["margin"] == pytest.approx(0.4)


@pytest.mark.skipif(not os.path.exists(S.MODEL_PATH), reason="no trained model committed")
def test_committed_model_is_well_formed():
    m = S.read_model(S.MODEL_PATH)
    _, total = OM.layout(m["sizes"])
    assert m["sizes"] == S.SIZES and len(m["params"]) == total
    assert np.all(np.isfinite(m["params"])) and 0 <= m["margin"] < 2
    assert m["limit"] == pytest.approx(S.LIMIT)


@pytest.mark.skipif(not os.path.exists(S.TIMES_PATH), reason="no timing table committed")
def test_committed_times_cover_every_supported_length():
    t = S.read_times(S.TIMES_PATH)
    assert set(S.timed_lengths()) <= set(t)
    assert all(v > 0 for v in t.values())
