"""CPU pins of the learned cache's oracle (oracle/mlp.py): the BF16 rounding, the frequency encoding,
the backward pass (central finite differences of the FP64 network) and the Adam step (closed forms of
its first steps)."""
import numpy as np

import mlp as M


def test_bf16_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -20, -2.5, 3.0e38,
                  1.0 + 2.0 ** -9], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7, -2.5, 2.9980e38, 1.0], np.float64)
    got = M.bf16(x).astype(np.float64)
    assert np.allclose(got[:5], want[:5], rtol=0, atol=0)
    assert got[6] == 1.0
    # 3e38 = 1.7633 * 2^127: mantissa step 2^120 -> the nearest bf16 value
    assert abs(got[5] - 3.0e38) <= 2.0 ** 119
    # every output is exactly representable in bf16 (low 16 bits clear)
    assert not np.any(M.bf16(np.random.default_rng(0).normal(size=1000)).view(np.uint32) & 0xFFFF)


def test_encoding_closed_forms():
    r = np.zeros(4, M.RECORD_DTYPE)
    shift = 8
    r["pos"][1] = [0.5 * 2.0 ** shift, 0.0, 0.0]          # band 0 of x at u = 0.5: sin 1, cos 0
    r["pos"][2] = [0.5 * 2.0 ** (shift - 11), 0.0, 0.0]   # band 11 of x at u = 0.5
    r["pos"][3] = [3.0, -7.25, 1000.5]
    r["n"][:] = [0.0, 1.0, 0.0]
    r["view"][:] = [0.6, 0.0, -0.8]
    X = M.encode(r, shift)
    assert X.shape == (4, 96)
    # pos 0: every sin 0 and every cos 1
    assert np.all(X[0, 0:72:2] == 0.0) and np.all(X[0, 1:72:2] == 1.0)
    assert abs(X[1, 0] - 1.0) < 1e-15 and abs(X[1, 1]) < 1e-15
    assert abs(X[2, 22] - 1.0) < 1e-15 and abs(X[2, 23]) < 1e-15
    # band k of coordinate a has period 2^(shift - k + 1)
    for a, k in [(0, 3), (1, 7), (2, 11)]:
        q = r.copy()
        q["pos"][3, a] += 2.0 ** (shift - k + 1)
        Y = M.encode(q, shift)
        assert abs(Y[3, 24 * a + 2 * k] - X[3, 24 * a + 2 * k]) < 1e-12
    assert np.all(X[:, 72:75] == [0.0, 1.0, 0.0]) and np.allclose(X[:, 75:78], [0.6, 0.0, -0.8])
    assert np.all(X[:, 78:] == 0.0)
    assert np.allclose(X[:, 0:72:2] ** 2 + X[:, 1:72:2] ** 2, 1.0)


def test_backward_matches_central_differences():
    params = M.init_params(3).astype(np.float64)
    recs = M.random_records(48, 7)
    loss, g, denom = M.loss_grad(params, recs, 8, 0.01, emulate=False)
    rng = np.random.default_rng(11)
    idx = np.concatenate([rng.choice(np.arange(M.w_off(l), M.b_off(l) + M.OUT_DIMS[l]), 12, replace=False)
                          for l in range(6)])
    # input rows 78..95 of layer 0 see zero features: their gradient is exactly zero
    assert np.all(g[M.w_off(0) + 78 * 64:M.b_off(0)] == 0.0)
    h = 1e-6
    bad = []
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        l1, _, _ = M.loss_grad(p1, recs, 8, 0.01, emulate=False, denom=denom)
        l2, _, _ = M.loss_grad(p2, recs, 8, 0.01, emulate=False, denom=denom)
        fd = (l1 - l2) / (2 * h)
        if abs(fd - g[i]) > 1e-6 * max(1.0, abs(g[i])) + 1e-9:
            bad.append((int(i), fd, g[i]))
    assert not bad, bad
    assert loss > 0


def test_backward_detects_a_planted_mistake():
    """A dropped ReLU mask in the backward pass is caught by the finite differences."""
    params = M.init_params(4).astype(np.float64)
    recs = M.random_records(32, 9)
    _, g, denom = M.loss_grad(params, recs, 8, 0.01, emulate=False)
    acts, y = M.forward(params, M.encode(recs, 8), emulate=False)
    # gradient of layer 3's weights without the mask of act[4] > 0
    layers = M.unpack(params)
    t = recs["rgb"].astype(np.float64)
    valid = (recs["tag"] != 0).astype(np.float64)[:, None]
    G = valid * (2.0 / 3.0) * (y - t) / denom[:, None] / len(recs)
    G = (G @ layers[5][0].T) * (acts[5] > 0)
    G = G @ layers[4][0].T   # mask dropped here
    wrong = acts[3].T @ G
    right = g[M.w_off(3):M.b_off(3)].reshape(64, 64)
    assert np.abs(wrong - right).max() > 1e-3 * np.abs(right).max()


def test_emulated_network_close_to_fp64():
    """BF16 operands move the gradient by a few percent (mostly ReLU units whose pre-activation changes
    sign under rounding; the error grows towards the input layer) but keep its direction."""
    params = M.init_params(5)
    recs = M.random_records(256, 13)
    l64, g64, _ = M.loss_grad(params, recs, 8, 0.01, emulate=False)
    lbf, gbf, _ = M.loss_grad(params, recs, 8, 0.01, emulate=True)
    assert abs(lbf - l64) < 0.02 * l64
    for l in range(6):
        a, b = M.w_off(l), M.b_off(l) + M.OUT_DIMS[l]
        cos = gbf[a:b] @ g64[a:b] / (np.linalg.norm(gbf[a:b]) * np.linalg.norm(g64[a:b]))
        assert cos > 0.99, (l, cos)
        assert np.linalg.norm(gbf[a:b] - g64[a:b]) < 0.15 * np.linalg.norm(g64[a:b])


def test_adam_first_steps_closed_form():
    rng = np.random.default_rng(2)
    w0 = rng.normal(size=1000).astype(np.float32)
    g = rng.normal(size=1000).astype(np.float32)
    m0 = np.zeros(1000, np.float32)
    lr, b1, b2, eps = 1e-3, 0.9, 0.99, 1e-8
    w1, m1, v1 = M.adam(w0, m0, m0, g, 1, lr, b1, b2, eps)
    # step 1: m_hat = g, v_hat = g^2 -> w1 = w0 - lr * g / (|g| + eps)
    want = w0.astype(np.float64) - lr * g / (np.abs(g) + eps)
    assert np.allclose(w1, want, rtol=1e-6, atol=1e-9)
    # step 2 with the same gradient: m_hat = g, v_hat = g^2 again
    w2, _, _ = M.adam(w1, m1, v1, g, 2, lr, b1, b2, eps)
    want2 = w1.astype(np.float64) - lr * g / (np.abs(g) + eps)
    assert np.allclose(w2, want2, rtol=1e-6, atol=1e-9)
    # zero gradient from zero moments leaves the weights unchanged
    w3, m3, v3 = M.adam(w0, m0, m0, np.zeros_like(g), 1, lr, b1, b2, eps)
    assert np.array_equal(w3, w0) and not m3.any() and not v3.any()


def test_weight_ema_closed_form():
    """The render weights' exponential mean: from we_0 toward a fixed w, we_k = w + ema^k (we_0 - w); ema 0
    is w itself; a constant sequence is a fixed point."""
    rng = np.random.default_rng(3)
    we0 = rng.normal(size=500).astype(np.float32)
    w = rng.normal(size=500).astype(np.float32)
    ema = 0.9
    we = we0
    for k in range(1, 6):
        we = M.weight_ema(we, w, ema)
        want = w.astype(np.float64) + (np.float64(np.float32(ema)) ** k) * (we0.astype(np.float64) - w)
        assert np.allclose(we, want, rtol=1e-5, atol=1e-6), k
    assert np.array_equal(M.weight_ema(we0, w, 0.0), w)
    assert np.allclose(M.weight_ema(w, w, ema), w, rtol=4e-7, atol=0)   # FP32 rounding of the two products only
