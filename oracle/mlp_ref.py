"""CPU restatement of the learned radius-proposal MLP (28 -> 64 -> 64 -> 1, ReLU, mean squared error, Adam).

TEST INFRASTRUCTURE ONLY: tests import this module to check the library's BF16-MFMA forward, backward and
Adam kernels. Nothing in the product path imports it.

The model has no counterpart in the reference (which always starts a query's search at its own cell,
GridKNNCuda.cu:2409-2474), so nothing here follows a reference line. Pinning:
  - `forward_backward(..., rounding=None)` is plain FP64 back-propagation; `tests/test_oracle_mlp.py` checks
    it against central finite differences of `loss()` on every parameter block (a wrong transpose, mask or
    sign fails that check);
  - with `rounding=bf16` the same code rounds exactly the values the device kernels round to bfloat16
    (every matrix-core operand) and computes all else in FP64, which the GPU tests compare against the
    library's FP32-accumulated results within a stated tolerance;
  - `adam()` is checked against a closed form of the first step (update = -lr * sign(g) up to eps).

Parameter vector layout (FP32): W1[32][64] (input-major, rows 28..31 zero padding), b1[64], W2[64][64],
b2[64], W3[64], b3.
"""
import numpy as np

NF, NIN, NH = 28, 32, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + NIN * NH
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NH
NPARAM = OFF_B3 + 1


def bf16(a):
    """Round float32 values to bfloat16 (round to nearest, ties to even), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(p):
    p = np.asarray(p)
    return dict(W1=p[OFF_W1:OFF_B1].reshape(NIN, NH), b1=p[OFF_B1:OFF_W2], W2=p[OFF_W2:OFF_B2].reshape(NH, NH),
                b2=p[OFF_B2:OFF_W3], W3=p[OFF_W3:OFF_B3], b3=p[OFF_B3])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([np.ravel(W1), np.ravel(b1), np.ravel(W2), np.ravel(b2), np.ravel(W3),
                           np.atleast_1d(b3)])


def pad_features(X):
    X = np.asarray(X)
    out = np.zeros((X.shape[0], NIN), X.dtype)
    out[:, :NF] = X
    return out


def _ident(a):
    return np.asarray(a, np.float64)


def _r(rounding, a):
    """Apply the operand rounding; FP32 intermediates are formed first where the device forms them."""
    if rounding is None:
        return np.asarray(a, np.float64)
    return rounding(np.asarray(a, np.float32)).astype(np.float64)


def forward_backward(params, X, T, rounding=None):
    """Returns (y, loss, grad) for features X (n x 28), targets T (n), batch-mean squared error.

    With rounding=None everything is FP64. With rounding=bf16 the operands the device feeds to its matrix
    cores are rounded: X, W1, W2, W3, the activations H1, H2 and the back-propagated dH2, dH1.
    """
    P = unpack(np.asarray(params, np.float64 if rounding is None else np.float32))
    X = pad_features(X)
    n = X.shape[0]
    f32 = (lambda a: np.asarray(a, np.float64)) if rounding is None else (lambda a: np.asarray(a, np.float32))
    Xb = _r(rounding, X)
    W1b, W2b, W3b = _r(rounding, P["W1"]), _r(rounding, P["W2"]), _r(rounding, P["W3"])
    b1, b2, b3 = (np.asarray(P[k], np.float64) for k in ("b1", "b2", "b3"))
    Z1 = f32(f32(Xb @ W1b) + f32(b1))
    H1 = _r(rounding, np.maximum(Z1, 0.0))
    Z2 = f32(f32(H1 @ W2b) + f32(b2))
    H2 = _r(rounding, np.maximum(Z2, 0.0))
    y = f32(f32(H2 @ W3b) + f32(b3))
    e = f32(y - np.asarray(T, np.float64))
    loss = float(np.sum(np.asarray(e, np.float64) ** 2) / n)
    dy = f32(f32(2.0 * e) * f32(1.0 / n)) if rounding is not None else 2.0 * e / n
    gW3 = H2.T @ dy
    gb3 = np.sum(dy)
    dH2 = _r(rounding, np.where(H2 > 0, f32(np.outer(dy, W3b)), 0.0))
    gb2 = dH2.sum(axis=0)
    gW2 = H1.T @ dH2
    dH1 = _r(rounding, np.where(H1 > 0, f32(dH2 @ W2b.T), 0.0))
    gb1 = dH1.sum(axis=0)
    gW1 = Xb.T @ dH1
    gW1[NF:, :] = 0.0
    grad = pack(gW1, gb1, gW2, gb2, gW3, gb3)
    return np.asarray(y, np.float64), loss, grad


def loss(params, X, T):
    """FP64 loss (finite-difference checks)."""
    return forward_backward(params, X, T, rounding=None)[1]


def adam(params, m, v, grad, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step in float32, as the device applies it. Returns (params, m, v)."""
    f = np.float32
    p = np.asarray(params, f)
    g = np.asarray(grad, f)
    m = f(beta1) * np.asarray(m, f) + (f(1.0) - f(beta1)) * g
    v = f(beta2) * np.asarray(v, f) + (f(1.0) - f(beta2)) * (g * g)
    mh = m / (f(1.0) - f(beta1) ** f(step))
    vh = v / (f(1.0) - f(beta2) ** f(step))
    return p - f(lr) * mh / (np.sqrt(vh) + f(eps)), m, v


def init_params(seed):
    """He-uniform draw used by tests that need a fresh parameter vector (any draw works for parity)."""
    rng = np.random.default_rng(seed)
    W1 = np.zeros((NIN, NH))
    W1[:NF] = rng.uniform(-1, 1, (NF, NH)) * np.sqrt(6.0 / NF)
    W2 = rng.uniform(-1, 1, (NH, NH)) * np.sqrt(6.0 / NH)
    W3 = rng.uniform(-1, 1, NH) * np.sqrt(3.0 / NH)
    b1 = rng.uniform(-0.1, 0.1, NH)
    b2 = rng.uniform(-0.1, 0.1, NH)
    return pack(W1, b1, W2, b2, W3, 0.25).astype(np.float32)


def synthetic_batch(n, seed):
    """Feature rows shaped like real ones: 27 log-normalised counts and log2(k)/8, plus smooth targets."""
    rng = np.random.default_rng(seed)
    counts = rng.poisson(rng.uniform(0.0, 40.0, (n, 1)), (n, 27))
    X = np.empty((n, NF), np.float32)
    X[:, :27] = np.log2(1.0 + counts) / 16.0
    X[:, 27] = np.log2(rng.choice([16, 40, 64], n)) / 8.0
    T = (1.5 - 0.8 * X[:, 13] * 16.0 / 6.0 + 0.3 * X[:, 27] + rng.normal(0, 0.05, n)).astype(np.float32)
    return X, T
