"""Learned predictor: MLP kernels against the FP64 oracle, training, and the
judged learned pass inside the AM2 scheme."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests import wake_harness as H

pytestmark = pytest.mark.gpu


def _ctx():
    from volcanor_amd import vlm
    return vlm.Context(0)


def test_forward_and_gradient_match_fp64_oracle():
    ctx = _ctx()
    n = ctx.mlp_num_params()
    assert n == 45 * 64 + 64 + 2 * (64 * 64 + 64) + 3 * 64 + 3
    rng = np.random.default_rng(1)
    p = (rng.normal(size=n) * 0.15).astype(np.float32)
    ctx.mlp_set_params(p)
    assert np.array_equal(ctx.mlp_get_params(), p)
    x = rng.normal(size=(300, 45)).astype(np.float32)
    t = rng.normal(size=(300, 3)).astype(np.float32)
    y = ctx.mlp_forward(x)
    yo = orc.mlp_forward(p.astype(np.float64), x.astype(np.float64))
    np.testing.assert_allclose(y, yo, rtol=2e-4, atol=2e-5)
    loss, g = ctx.mlp_loss_grad(x, t)
    lo, go = orc.mlp_loss_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    assert loss == pytest.approx(lo, rel=1e-4)
    scale = np.max(np.abs(go))
    np.testing.assert_allclose(g, go, rtol=1e-3, atol=2e-5 * scale)


def test_training_reduces_loss_on_smooth_target():
    ctx = _ctx()
    rng = np.random.default_rng(3)
    x = rng.normal(size=(2048, 45)).astype(np.float32)
    w = rng.normal(size=(45, 3)) / 10
    t = np.tanh(x @ w).astype(np.float32)
    first = ctx.mlp_train_step(x, t, lr=3e-3)
    for _ in range(150):
        last = ctx.mlp_train_step(x, t, lr=3e-3)
    assert last < 0.2 * first


def test_learned_pass_in_am2_tracks_full_pass():
    """Free-wake rotor with the learned pass 2: every step is either judged
    within tolerance on the sampled targets or recomputed; the wake stays
    close to the plain (full pass 2) wake."""
    case = H.Case(nb=2, nc=1, ns=8, nNwake=30, nFwake=4, nt=40, scheme=3, omega=30.0,
                  vel=(0.0, 0.0, 1.0))
    plain, learned = _ctx(), _ctx()
    learned.set_predictor(1, tol=1e-3, sample_stride=4, warmup=6, train_iters=8, lref=0.25)
    ia, ib = H.setup(plain, case), H.setup(learned, case)
    for it in range(1, case.nt + 1):
        H.step(plain, case, ia, it)
        H.step(learned, case, ib, it)
    st = learned.predictor_stats()
    assert st["full"] + st["learned"] == case.nt - 1
    assert st["full"] >= 6
    a = plain.rotor_get_wake(ia[0], case.dims)["nodes"]
    b = learned.rotor_get_wake(ib[0], case.dims)["nodes"]
    assert np.max(np.abs(a - b)) < 1e-3 * np.max(np.abs(a))
