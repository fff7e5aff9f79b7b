"""Decode-cost model: the double-precision checker pinned by finite
differences and closed forms, the device's forward/backward/Adam kernels
against it, and what the trained model gains over the compressed-length
heuristic on decode times the lane kernel records."""
import numpy as np
import pytest

import signinum_amd as S
from signinum_amd import synth
from tests import oracle_ref as O

N_PARAMS = 1281
W1, B1, W2, B2, W3, B3 = 0, 160, 192, 1216, 1248, 1280


def _random_params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.normal(0, scale, N_PARAMS)


def _data(seed, n):
    rng = np.random.default_rng(seed)
    X = rng.random((n, 5)).astype(np.float32)
    Y = (np.sin(3 * X[:, 0]) + X[:, 1] * X[:, 4] - 0.5 * X[:, 2] + 0.3 * X[:, 3]).astype(np.float32)
    return X, Y


def test_oracle_gradient_matches_finite_differences(oracle):
    p = _random_params(3)
    X, Y = _data(4, 40)
    _loss, g = O.mlp_loss_grad(p, X, Y)
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(N_PARAMS, 60, replace=False), [B1, B2, W3, B3]])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        num = (O.mlp_loss(pp, X, Y) - O.mlp_loss(pm, X, Y)) / (2 * h)
        assert abs(num - g[k]) <= 1e-6 + 1e-5 * abs(num), (k, num, g[k])


def test_oracle_forward_closed_forms(oracle):
    # only the output bias: every prediction is b3, loss is the mean of 0.5 (b3 - y)^2
    p = np.zeros(N_PARAMS)
    p[B3] = 0.75
    X, Y = _data(6, 30)
    assert np.allclose(O.mlp_predict(p, X), 0.75)
    assert np.isclose(O.mlp_loss(p, X, Y), np.mean(0.5 * (0.75 - Y.astype(np.float64)) ** 2))
    # one path: h1_0 = relu(x_2), h2_0 = relu(2 h1_0), f = 3 h2_0 - 1
    p = np.zeros(N_PARAMS)
    p[W1 + 0 * 5 + 2] = 1.0
    p[W2 + 0 * 32 + 0] = 2.0
    p[W3 + 0] = 3.0
    p[B3] = -1.0
    x = np.array([[0.1, 0.2, 0.3, 0.4, 0.5], [0.1, 0.2, -0.3, 0.4, 0.5]], np.float32)
    assert np.allclose(O.mlp_predict(p, x), [6 * np.float32(0.3) - 1, -1.0])


@pytest.mark.gpu
def test_device_gradient_matches_oracle(session, oracle):
    model = S.CostModel(session, seed=11)
    p = model.params
    X, Y = _data(12, 3000)
    g_dev = model.gradient(X, Y)
    _loss, g_ref = O.mlp_loss_grad(p.astype(np.float64), X, Y)
    scale = np.abs(g_ref).max()
    assert np.allclose(g_dev, g_ref, rtol=1e-3, atol=1e-5 * scale), np.abs(g_dev - g_ref).max()
    model.close()


@pytest.mark.gpu
def test_device_training_reduces_loss_and_predicts_like_oracle(session, oracle):
    model = S.CostModel(session, seed=13)
    X, Y = _data(14, 4000)
    loss0 = O.mlp_loss(model.params.astype(np.float64), X, Y)
    loss = model.train(X, Y, epochs=400, lr=1e-2)
    assert loss < loss0 / 10, (loss0, loss)
    p = model.params.astype(np.float64)
    assert np.isclose(O.mlp_loss(p, X, Y), loss, rtol=1e-3)
    Xt, _ = _data(15, 64)
    assert np.allclose(model.predict(Xt), O.mlp_predict(p, Xt), rtol=1e-4, atol=1e-4)
    model.close()


@pytest.mark.gpu
def test_model_gain_over_length_heuristic(session, oracle):
    """Trained on lane-kernel timings, the model predicts held-out decode
    times better than the compressed-length heuristic fitted to the same
    data; the kernel time of the orders built on each is reported."""
    streams, pix, _cls = synth.lossless_batch_with_pixels(768, size=256)
    plan = S.TilePlan(session, streams, S.PixelFormat.RGB8)
    feats, cyc = plan.time_code_blocks(stride=3, max_samples=20000)
    assert len(cyc) > 4000 and (cyc > 0).all()
    y = np.log2(cyc).astype(np.float32)
    rng = np.random.default_rng(16)
    order = rng.permutation(len(y))
    tr, te = order[: len(y) * 3 // 4], order[len(y) * 3 // 4:]
    mean, std = y[tr].mean(), y[tr].std()
    model = S.CostModel(session, seed=17)
    model.train(feats[tr], (y[tr] - mean) / std, epochs=1500, lr=3e-3)
    pred = model.predict(feats[te]) * std + mean
    mse_model = float(np.mean((pred - y[te]) ** 2))
    # heuristic: log2(coded bytes + samples / 4), affine fit on the training split
    h = np.log2(feats[:, 0] * 4096 + feats[:, 4] * 4096 / 4)
    A = np.stack([h[tr], np.ones(len(tr))], 1)
    coef, *_ = np.linalg.lstsq(A, y[tr], rcond=None)
    mse_heur = float(np.mean((h[te] * coef[0] + coef[1] - y[te]) ** 2))
    print(f"held-out MSE of log2(cycles): model {mse_model:.4f}, length heuristic {mse_heur:.4f}")
    assert mse_model < mse_heur

    def lane_ms():
        ev = [S.HipEvent(), S.HipEvent()]
        plan.decode_stage(0)
        ev[0].record(session)
        for _ in range(5):
            plan.decode_stage(0)
        ev[1].record(session)
        session.synchronize()
        return ev[0].elapsed_ms(ev[1]) / 5

    plan.order_code_blocks(plan.ORDER_GEOMETRY_COST)
    plan.autotune(reps=2)
    t = {}
    for name, o, m in (("input order", plan.ORDER_INPUT, None),
                       ("length heuristic", plan.ORDER_LENGTH, None),
                       ("model", plan.ORDER_MODEL, model),
                       ("shape + tile length heuristic", plan.ORDER_GEOMETRY_COST, None),
                       ("shape + tile model", plan.ORDER_GEOMETRY_MODEL, model)):
        plan.order_code_blocks(o, m)
        t[name] = lane_ms()
        plan.decode(S.MODE_GRAPH)
        session.synchronize()
        for i in range(0, len(pix), 97):
            assert np.array_equal(plan.surfaces[i].to_numpy(), pix[i]), (name, i)
    print("HT lane stage ms by code-block order:", {k: round(v, 4) for k, v in t.items()})
    model.close()
    plan.close()
