"""The roundoff MLP's HIP kernels and the learned FFT-length selector on the GPU.

The kernels (ll_mlp_*) are checked against the float64 numpy oracle
(oracle/mlp.py, itself pinned by finite differences and closed forms in
test_selector.py): forward values, the loss gradient, and an Adam step.  The
selector is checked for what it promises: lengths it picks are supported and
inside its domain (never longer than the reference policy's without a timing
table), the fallback moves up, and on the seeded evaluation exponents it
saves projected and batched time with every residue equal to the
reference-length run and to GMP.
"""
import os

import numpy as np
import pytest

import cudalucas_amd as L
import oracle as O
from oracle import mlp as OM
from cudalucas_amd import selector as S
from cudalucas_amd import workloads as W
from conftest import announce

pytestmark = pytest.mark.gpu

SIZES = [7, 64, 64, 64, 1]


def random_net(seed, sizes=SIZES):
    rng = np.random.default_rng(seed)
    _, total = OM.layout(sizes)
    return (rng.standard_normal(total) * 0.3).astype(np.float32)


def test_mlp_forward_matches_oracle():
    params = random_net(1)
    x = np.random.default_rng(2).standard_normal((300, 7)).astype(np.float32)
    with L.MLP(SIZES) as m:
        assert m.count == len(params)
        m.set_params(params)
        assert (m.get_params() == params).all()
        y = m.forward(x)
    yo, _ = OM.forward(params.astype(np.float64), SIZES, x.astype(np.float64))
    # FP32 through 4 layers: a few ulps of the activations' scale
    assert np.max(np.abs(y - yo)) < 2e-5 * max(1.0, np.max(np.abs(yo)))


@pytest.mark.parametrize("sizes,batch", [(SIZES, 256), ([7, 64, 64, 64, 1], 33), ([3, 5, 1], 7), ([7, 1], 64)])
def test_mlp_gradient_matches_oracle(sizes, batch):
    params = random_net(3, sizes)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((batch, sizes[0])).astype(np.float32)
    t = rng.standard_normal(batch).astype(np.float32)
    with L.MLP(sizes) as m:
        m.set_params(params)
        loss, g = m.gradient(x, t)
        assert (m.get_params() == params).all()  # gradient does not update
    lo, go = OM.loss_and_grad(params.astype(np.float64), sizes, x.astype(np.float64), t.astype(np.float64))
    assert loss == pytest.approx(lo, rel=1e-5)
    # componentwise, relative to the gradient's scale (FP32 sums over the batch)
    assert np.max(np.abs(g - go)) < 1e-5 * np.max(np.abs(go)) + 1e-7


def test_mlp_adam_steps_match_oracle():
    sizes = [7, 16, 16, 1]
    params = random_net(5, sizes)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((64, 7)).astype(np.float32)
    t = rng.standard_normal(64).astype(np.float32)
    lr = 1e-2
    with L.MLP(sizes) as m:
        m.set_params(params)
        m.train(x, t, batch=64, epochs=3, lr=lr, seed=9)  # one batch per epoch: order irrelevant to the sum
        got = m.get_params().astype(np.float64)
    p = params.astype(np.float64)
    mo = np.zeros_like(p)
    vo = np.zeros_like(p)
    gmin = np.full_like(p, np.inf)
    for step in range(1, 4):
        _, g = OM.loss_and_grad(p, sizes, x.astype(np.float64), t.astype(np.float64))
        gmin = np.minimum(gmin, np.abs(g) / np.max(np.abs(g)))
        p, mo, vo = OM.adam_step(p, g, mo, vo, step, lr)
    # Adam divides by |g|: where a gradient component is tiny, FP32 rounding
    # of the batch sum moves the step by up to lr; elsewhere it must agree
    well = gmin > 1e-3
    assert well.mean() > 0.9
    assert np.max(np.abs(got - p)[well]) < 5e-5
    assert np.max(np.abs(got - p)) <= 3 * lr * 1.001
    assert np.max(np.abs(got - params)) > 0.5 * lr  # the steps were taken


def test_mlp_training_fits_smooth_function():
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, (4096, 7)).astype(np.float32)
    t = (np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2] - 0.5 * x[:, 3] ** 2).astype(np.float32)
    with L.MLP(SIZES) as m:
        losses = m.train(x, t, batch=128, epochs=60, lr=3e-3, seed=1)
        y = m.forward(x)
    assert losses[-1] < 0.05 * losses[0]
    assert np.sqrt(np.mean((y - t) ** 2)) < 0.06


def test_mlp_rejects_bad_shapes():
    with pytest.raises(L.LLError):
        L.MLP([7, 200, 1])      # wider than the kernels' limit
    with pytest.raises(L.LLError):
        L.MLP([7, 64, 2])       # the loss takes one output


# ------------------------------------------------------------ selector

needs_model = pytest.mark.skipif(not (os.path.exists(S.MODEL_PATH) and os.path.exists(S.TIMES_PATH)),
                                 reason="no trained model / timing table committed")


@needs_model
def test_selector_predictions_follow_the_model_file():
    model = S.read_model(S.MODEL_PATH)
    params = np.asarray(model["params"])
    with S.Selector(S.MODEL_PATH) as sel:
        for p in (100_003, 1_000_003, 19_999_999, 77_005_277):
            ns = S.candidates(p)
            got = np.log2(sel.predict(p, ns))
            x = (S.feature_matrix([p] * len(ns), ns) - np.asarray(model["feature_mean"])) / np.asarray(model["feature_scale"])
            y, _ = OM.forward(params, model["sizes"], x)
            want = y * model["target"][1] + model["target"][0]
            want += model.get("growth_scale", 1.0) * np.maximum(0.0, S.growth(ns, p - 2, model["window"]))
            want += model["margin"]
            assert np.max(np.abs(got - want)) < 1e-3, p


@needs_model
def test_selector_choices_are_supported_and_move_up_on_fallback():
    times = S.read_times(S.TIMES_PATH)
    with S.Selector(S.MODEL_PATH, times=times) as sel:
        for p in W.log_uniform_primes(99, 40, 100_000, 120_000_000):
            nref = L.choose_fft_length(p)[0]
            n, e = sel.choose(p)
            assert L.fft_supported(n) and n <= 2 * nref and n in times
            assert n == nref or e <= S.LIMIT
            up, _ = sel.choose(p, above=n)
            assert up > n and L.fft_supported(up)


@needs_model
def test_selector_without_times_never_exceeds_reference_length():
    with S.Selector(S.MODEL_PATH) as sel:
        for p in W.log_uniform_primes(98, 40, 100_000, 120_000_000):
            nref = L.choose_fft_length(p)[0]
            n, e = sel.choose(p)
            assert n <= nref and (n == nref or e <= S.LIMIT)


@needs_model
def test_selector_full_evaluation_all_residues_vs_gmp(gmp_fixtures):
    """All 512 seeded evaluation primes in [1e5, 2e7]: the selected lengths
    save projected one-at-a-time time and measured batched time against the
    reference policy, every residue after the round-off test equals the
    reference-length run's, there is no unresolved fallback, and every
    Res64 after the round-off test equals GMP's (committed fixtures,
    tests/golden/make_gmp_fixtures.py)."""
    import json

    from cudalucas_amd import tuner as T

    primes = W.log_uniform_primes(W.SELECTOR_EVAL_SEED, W.SELECTOR_EVAL_COUNT, *W.SELECTOR_EVAL_RANGE)
    missing = [p for p in primes if (p, S.CHECK_ITERS) not in gmp_fixtures]
    assert not missing, "missing GMP fixtures for %d exponents" % len(missing)
    times = S.read_times(S.TIMES_PATH)
    btimes = S.read_times(S.BATCH_TIMES_PATH)
    sched = T.load(T.BATCH_TABLE)
    announce("selector full evaluation on %d exponents" % len(primes))
    rep = S.evaluate(S.MODEL_PATH, primes, times, batch_times=btimes, schedules=sched, log=print)
    out = os.environ.get("LL_SELECTOR_REPORT")
    if out:
        with open(out, "w") as f:
            json.dump(rep, f, indent=1)
    print({k: v for k, v in rep.items() if k != "per_exponent"})
    for form in ("timed", "batched"):
        assert rep[form]["residue_mismatches_vs_reference_length"] == []
        assert rep[form]["max_roundoff"] <= S.LIMIT
    assert rep["reference_length_check_failures"] == 0
    assert rep["time_saved_fraction_plain"] > 0.0
    # the batched engine at the selected lengths against the plain
    # per-exponent path at the reference policy's lengths
    assert rep["batched_selected_vs_plain_reference"] > 2.0
    got = {row["p"]: row["res64"] for row in rep["per_exponent"]}
    bad = [p for p in primes if "0x" + got[p] != gmp_fixtures[(p, S.CHECK_ITERS)]["res64"]]
    print("GMP Res64 checked for %d exponents, %d mismatches" % (len(primes), len(bad)))
    assert bad == []
