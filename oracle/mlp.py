"""fp64 reference of the (W,P)-selector MLP (TEST INFRASTRUCTURE ONLY).

64 -> 128 (ReLU) -> 128 (ReLU) -> 3, weights used as bf16 (round to nearest even), loss =
mean squared error over all outputs.  Parameters are one flat fp32 vector in the layout
W1[128][64] b1[128] W2[128][128] b2[128] W3[3][128] b3[3].  The gradient is with respect to
the (bf16-rounded) weights, which is what the device kernels compute.
"""
import numpy as np

SHAPES = [("W1", (128, 64)), ("b1", (128,)), ("W2", (128, 128)), ("b2", (128,)),
          ("W3", (3, 128)), ("b3", (3,))]
COUNT = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16_round(a: np.ndarray) -> np.ndarray:
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (r.astype(np.uint32) << 16).view(np.float32)


def unpack(params: np.ndarray) -> dict:
    out, off = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = params[off:off + k].reshape(shape)
        off += k
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def forward(params: np.ndarray, X: np.ndarray):
    p = unpack(np.asarray(params, dtype=np.float64))
    W1, W2, W3 = (bf16_round(p[k]).astype(np.float64) for k in ("W1", "W2", "W3"))
    X = np.asarray(X, dtype=np.float64)
    a1 = X @ W1.T + p["b1"]
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2.T + p["b2"]
    h2 = np.maximum(a2, 0)
    y = h2 @ W3.T + p["b3"]
    return y, (X, a1, h1, a2, h2, W1, W2, W3)


def loss_grad(params: np.ndarray, X: np.ndarray, Y: np.ndarray):
    y, (X, a1, h1, a2, h2, W1, W2, W3) = forward(params, X)
    Y = np.asarray(Y, dtype=np.float64)
    N = y.size
    e = y - Y
    loss = float((e * e).sum() / N)
    dy = 2 * e / N
    g = {"W3": dy.T @ h2, "b3": dy.sum(0)}
    dh2 = (dy @ W3) * (a2 > 0)
    g["W2"], g["b2"] = dh2.T @ h1, dh2.sum(0)
    dh1 = (dh2 @ W2) * (a1 > 0)
    g["W1"], g["b1"] = dh1.T @ X, dh1.sum(0)
    return loss, pack(g)
