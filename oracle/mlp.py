"""CPU oracle of the learned GI radiance cache's network — TEST INFRASTRUCTURE ONLY.

A numpy restatement of the MLP that libvftrace.so trains and evaluates on the matrix cores
(include/vf_gicache.h): frequency encoding, five hidden layers of 64 with ReLU, three outputs, the
relative L2 loss, its gradient and the Adam step.  The reference has no network (its radiance cache
is a spatial hash, shaders/common.glsl:2150-2230), so this module restates the library's own
definition.  It is used by tests/ only, as the checker:

* `grad(..., emulate=False)`: the plain FP64 network.  Pinned by central finite differences of the
  loss (tests/test_mlp_oracle.py).
* `grad(..., emulate=True)`: the same arithmetic with the kernel's BF16 rounding points (inputs,
  weights, activations and backpropagated gradients rounded to BF16 round-to-nearest-even; sums in
  FP64 where the kernel accumulates in FP32).  GPU results must agree with it to FP32 accumulation
  error.
* `adam(...)`: the FP32 update in the kernel's operation order (bit-for-bit comparable).
"""
import numpy as np

IN_DIMS = (96, 64, 64, 64, 64, 64)
OUT_DIMS = (64, 64, 64, 64, 64, 3)
N_PARAMS = 23043
N_INPUTS = 96
N_FEATURES = 78
N_BANDS = 12

RECORD_DTYPE = np.dtype([("pos", "<f4", (3,)), ("tag", "<u4"), ("n", "<f4", (3,)), ("pad0", "<f4"),
                         ("view", "<f4", (3,)), ("pad1", "<f4"), ("rgb", "<f4", (3,)), ("pad2", "<f4")])


def w_off(l):
    return 0 if l == 0 else 96 * 64 + 64 + (l - 1) * (64 * 64 + 64)


def b_off(l):
    return w_off(l) + IN_DIMS[l] * OUT_DIMS[l]


def unpack(params):
    """Parameter vector -> [(W_l [in, out], b_l [out])]."""
    p = np.asarray(params)
    assert p.size == N_PARAMS
    out = []
    for l in range(6):
        W = p[w_off(l):b_off(l)].reshape(IN_DIMS[l], OUT_DIMS[l])
        b = p[b_off(l):b_off(l) + OUT_DIMS[l]]
        out.append((W, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.reshape(-1), b]) for W, b in layers])


def bf16(x):
    """Round FP32 values to BF16 (round to nearest, ties to even); returns float32 holding them."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r.astype(np.uint32)).view(np.float32).reshape(a.shape)


def encode(recs, shift):
    """Input features (FP64, n x 96): feature 24a + 2k (+1) = sin (cos) of pi * pos[a] * 2^(k - shift),
    k = 0..11; 72..74 normal, 75..77 view, 78..95 zero.  pos * 2^(k - shift) is exact in FP32."""
    r = np.asarray(recs)
    n = r.shape[0]
    X = np.zeros((n, N_INPUTS), np.float64)
    for a in range(3):
        for k in range(N_BANDS):
            u = np.ldexp(r["pos"][:, a].astype(np.float32), k - shift).astype(np.float64)
            X[:, 24 * a + 2 * k] = np.sin(np.pi * u)
            X[:, 24 * a + 2 * k + 1] = np.cos(np.pi * u)
    X[:, 72:75] = r["n"]
    X[:, 75:78] = r["view"]
    return X


def forward(params, X, emulate):
    """Returns (acts, y): acts[l] = input of layer l (acts[0] = X), y = outputs (n x 3)."""
    layers = unpack(np.asarray(params, np.float64))
    a = bf16(X) if emulate else np.asarray(X, np.float64)
    acts = [a.astype(np.float64)]
    for l in range(6):
        W, b = layers[l]
        if emulate:
            W = bf16(W).astype(np.float64)
        z = acts[-1] @ W + b
        if l == 5:
            return acts, z
        h = np.maximum(z, 0.0)
        if emulate:
            h = bf16(h.astype(np.float32)).astype(np.float64)
        acts.append(h)
    raise AssertionError


def loss_grad(params, recs, shift, loss_eps, emulate, X=None, denom=None):
    """Mean relative L2 loss over the n records (invalid ones count in n but add nothing) and its
    gradient (FP64, parameter order).  The denominator mean(y)^2 + loss_eps is held constant (a
    stop-gradient), as in the kernel; `denom` overrides it (finite differences)."""
    r = np.asarray(recs)
    n = r.shape[0]
    if X is None:
        X = encode(r, shift)
    acts, y = forward(params, X, emulate)
    t = r["rgb"].astype(np.float64)
    valid = (r["tag"] != 0).astype(np.float64)[:, None]
    if denom is None:
        m = y.mean(axis=1)
        denom = m * m + loss_eps
    d = y - t
    loss = float(np.sum(valid * d * d / denom[:, None]) / 3.0 / n)
    gy = valid * (2.0 / 3.0) * d / denom[:, None] / n
    layers = unpack(np.asarray(params, np.float64))
    G = bf16(gy.astype(np.float32)).astype(np.float64) if emulate else gy
    grads = [None] * 6
    for l in range(5, -1, -1):
        W, _ = layers[l]
        if emulate:
            W = bf16(W).astype(np.float64)
        grads[l] = (acts[l].T @ G, G.sum(axis=0))
        if l == 0:
            break
        Gp = (G @ W.T) * (acts[l] > 0)
        G = bf16(Gp.astype(np.float32)).astype(np.float64) if emulate else Gp
    return loss, pack(grads), denom


def weight_ema(we, w, ema):
    """The render's weights after one step: ema * we + (1 - ema) * w in FP32, the kernel's order."""
    f = np.float32
    return (f(ema) * np.asarray(we, np.float32) + (f(1.0) - f(ema)) * np.asarray(w, np.float32)).astype(np.float32)


def adam(w, m, v, g, step, lr, b1, b2, eps):
    """One Adam step in FP32 with the kernel's operation order; returns (w, m, v)."""
    f = np.float32
    w, m, v, g = (np.asarray(x, np.float32) for x in (w, m, v, g))
    c1, c2 = f(1.0) - f(b1), f(1.0) - f(b2)
    # bias corrections: 1 - beta^t in double from the FP32 betas, then rounded to FP32
    bc1 = f(1.0 - np.float64(np.float32(b1)) ** step)
    bc2 = f(1.0 - np.float64(np.float32(b2)) ** step)
    m2 = f(b1) * m + c1 * g
    v2 = f(b2) * v + c2 * (g * g)
    mh = m2 / bc1
    vh = v2 / bc2
    w2 = w - f(lr) * mh / (np.sqrt(vh) + f(eps))
    return w2.astype(np.float32), m2.astype(np.float32), v2.astype(np.float32)


def init_params(seed):
    """He-uniform weights from a numpy generator (tests; the library has its own initialisation)."""
    rng = np.random.default_rng(seed)
    layers = []
    for l in range(6):
        fan = 78 if l == 0 else 64
        bound = np.sqrt(6.0 / fan) * (0.25 if l == 5 else 1.0)
        W = rng.uniform(-bound, bound, (IN_DIMS[l], OUT_DIMS[l]))
        if l == 0:
            W[78:] = 0.0
        b = rng.uniform(-0.1, 0.1, OUT_DIMS[l])
        layers.append((W, b))
    return pack(layers).astype(np.float32)


def random_records(n, seed, extent=256.0, valid_frac=0.9):
    """Synthetic records: positions in the world box, unit normals / view directions, RGB targets."""
    rng = np.random.default_rng(seed)
    r = np.zeros(n, RECORD_DTYPE)
    r["pos"] = rng.uniform(0.0, extent, (n, 3)).astype(np.float32)
    nv = rng.normal(size=(n, 3))
    r["n"] = (nv / np.linalg.norm(nv, axis=1, keepdims=True)).astype(np.float32)
    vv = rng.normal(size=(n, 3))
    r["view"] = (vv / np.linalg.norm(vv, axis=1, keepdims=True)).astype(np.float32)
    r["tag"] = (rng.uniform(size=n) < valid_frac).astype(np.uint32)
    r["rgb"] = rng.uniform(0.0, 2.0, (n, 3)).astype(np.float32)
    return r
