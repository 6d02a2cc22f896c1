"""HIP surrogate network (bf16 MFMA, FP32 accumulation) against its CPU oracle.

Tolerances (relative L2): forward 1e-3 against the oracle with the same
bf16 rounding points; gradients 2e-2 per layer (a different FP32/FP64 sum
can move an intermediate value across a bf16 rounding boundary, which then
propagates); the loss 1e-4.  The oracle's own backward pass is pinned by
finite differences in tests/test_surrogate_oracle.py.
"""
import os

import numpy as np
import pytest

from oracle import surrogate_oracle as so
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

NB, CH = 6, 32


def _case(seed=0, B=2, H=16, W=16):
    rng = np.random.default_rng(seed)
    theta = so.init_params(seed, NB, CH).astype(np.float64)
    theta += 0.02 * rng.standard_normal(theta.size)     # nonzero biases everywhere
    theta = theta.astype(np.float32)
    x = rng.standard_normal((B, H, W, 2)).astype(np.float32)
    t = (x[..., 0] + 0.1 * rng.standard_normal((B, H, W))).astype(np.float32)
    return theta, x, t


@pytest.fixture(scope="module")
def net():
    from multem_amd.surrogate import SurrogateNet
    n = SurrogateNet(NB, CH)
    yield n
    n.close()


def test_param_roundtrip(net):
    theta, _, _ = _case(1)
    assert net.n_params == so.n_params(NB, CH)
    net.set_params(theta)
    assert np.array_equal(net.get_params(), theta)


@pytest.mark.parametrize("shape", [(2, 16, 16), (1, 32, 48)])
def test_forward_matches_oracle(net, shape):
    theta, x, _ = _case(2, *shape)
    net.set_params(theta)
    y = net.forward(x)
    y_bf, _ = so.forward(theta, x, NB, CH, rounding="bf16")
    y64, _ = so.forward(theta, x, NB, CH, rounding="none")
    e_bf, e64 = rel_l2(y, y_bf), rel_l2(y, y64)
    print("forward %s: rel L2 vs bf16 oracle %.2e, vs FP64 %.2e" % (shape, e_bf, e64))
    assert e_bf < 1e-3
    assert e64 < 2e-2


def test_gradients_match_oracle(net):
    theta, x, t = _case(3)
    net.set_params(theta)
    loss, g = net.loss_grad(x, t)
    loss_ref, g_ref = so.loss_and_grad(theta, x, t, NB, CH, rounding="bf16")
    assert abs(loss - loss_ref) <= 1e-4 * loss_ref
    o = 0
    worst = 0.0
    for name, ws, bs in so.layer_shapes(NB, CH):
        nw, nb = int(np.prod(ws)), int(np.prod(bs))
        ew = rel_l2(g[o:o + nw], g_ref[o:o + nw])
        eb = rel_l2(g[o + nw:o + nw + nb], g_ref[o + nw:o + nw + nb])
        worst = max(worst, ew, eb)
        assert ew < 2e-2 and eb < 2e-2, (name, ew, eb)
        o += nw + nb
    print("gradients: worst per-layer rel L2 vs bf16 oracle %.2e" % worst)
    assert rel_l2(g, g_ref) < 1e-2


def test_training_reduces_loss_deterministically(net):
    rng = np.random.default_rng(4)
    N, H, W = 24, 16, 16
    x = rng.standard_normal((N, H, W, 2)).astype(np.float32)
    # target: a local, nonlinear function of both channels (learnable by 3x3 convs)
    sm = 0.25 * (np.roll(x[..., 1], 1, 1) + np.roll(x[..., 1], -1, 1) + np.roll(x[..., 1], 1, 2)
                 + np.roll(x[..., 1], -1, 2))
    t = (x[..., 0] + 0.5 * sm - 0.2 * np.maximum(x[..., 0], 0)).astype(np.float32)
    theta0 = so.init_params(9, NB, CH)
    net.set_params(theta0)
    l1 = net.train(x, t, batch=8, steps=150, lr=2e-3, seed=11)
    p1 = net.get_params()
    net.set_params(theta0)
    l2 = net.train(x, t, batch=8, steps=150, lr=2e-3, seed=11)
    p2 = net.get_params()
    print("training loss %.4f -> %.4f" % (l1[:5].mean(), l1[-10:].mean()))
    assert np.array_equal(l1, l2) and np.array_equal(p1, p2)
    assert l1[-10:].mean() < 0.3 * l1[:5].mean()


_DATA = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "multem_amd", "data")
WEIGHTS = os.path.join(_DATA, "surrogate_6x32.npz")
EVAL = os.path.join(_DATA, "surrogate_eval.npz")


def _rel_rms(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.sqrt(np.mean((a - b) ** 2, axis=(-2, -1)) / np.mean(b ** 2, axis=(-2, -1)))


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="no trained surrogate weights")
def test_trained_surrogate_gain():
    """Held-out pairs (never seen in training; the last 40 of each of the five
    data files: 80 untilted, 120 with illumination tilt up to 1 degree): the
    network's error against the 64-configuration image, next to that of the
    2-configuration image it starts from.

    Measured with the committed weights (the CPU oracle's bf16 forward pass
    gives the same figures): untilted mean ratio 0.51, network better on 79
    of 80; tilted mean ratio 0.73, better on 99 of 120.  The bars below sit
    above those figures as regression guards.  The 64-configuration image is
    not the converged one (its own phonon noise is about 1.5 %), so the
    network is measured against a converged estimate in
    surrogate.evaluate and in the bench."""
    from multem_amd.surrogate import Surrogate
    model = Surrogate.load(WEIGHTS)
    d = np.load(EVAL)
    pred = model.predict(d["adf_small"], d["vmap"])
    e_net = _rel_rms(pred, d["adf_big"])
    e_base = _rel_rms(d["adf_small"], d["adf_big"])
    tilted = d["tilted"].astype(bool)
    for name, sel, ratio, better in (("untilted", ~tilted, 0.6, 0.95), ("tilted", tilted, 0.8, 0.75)):
        en, eb = e_net[sel], e_base[sel]
        print("held-out %s pairs %d: rel RMS vs 64 configs: network %.4f (max %.4f), 2 configs %.4f (max %.4f); "
              "ratio %.3f, network better on %d" % (name, len(en), en.mean(), en.max(), eb.mean(), eb.max(),
                                                    en.mean() / eb.mean(), int(np.sum(en < eb))))
        assert len(en) > 0
        assert en.mean() < ratio * eb.mean(), name
        assert np.mean(en < eb) >= better, name


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="no trained surrogate weights")
def test_surrogate_gpu_time_ratio():
    from multem_amd import surrogate
    model = surrogate.Surrogate.load(WEIGHTS)
    ev = surrogate.evaluate(model, n_specimens=2, seed=2024)
    print("GPU time plain 64 configs / surrogate: %.1f; rel RMS surrogate %.4f, 2 configs %.4f"
          % (ev["gpu_time_ratio"], ev["rel_rms_surrogate_mean"], ev["rel_rms_small_mean"]))
    assert ev["gpu_time_ratio"] >= 8.0
    assert ev["rel_rms_surrogate_mean"] < ev["rel_rms_small_mean"]
