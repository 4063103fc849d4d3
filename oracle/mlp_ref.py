"""CPU oracle of the round-scheduler network (numpy, float64 by default).

TEST INFRASTRUCTURE ONLY: the GPU forward/backward/Adam kernels
(parallel-task1_amd/csrc/mlp.hip) are checked against these functions. The
network has no counterpart in the reference; its definition is the one the
learned schedule specifies: 5 inputs -> 64 ReLU -> 64 ReLU -> 3 mode logits +
4 chunk-bucket logits, loss = mean over rows of the two softmax
cross-entropies.

Parameter layout (row-major W[out][in]): W1 (64x5), b1, W2 (64x64), b2,
W3 (7x64), b3 -- the same order the library uses.

Pins: the analytic backward is checked against central finite differences in
tests/test_mlp_oracle.py (a dropped term, a transposed weight or a wrong
softmax sign fails it), and the forward against a direct per-row loop.
"""
import numpy as np

IN, HID, MODES, CHUNKS = 5, 64, 3, 4
OUT = MODES + CHUNKS
SHAPES = [("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (OUT, HID)), ("b3", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([d[name].ravel() for name, _ in SHAPES])


def forward(p, x):
    P = unpack(p)
    z1 = x @ P["W1"].T + P["b1"]
    h1 = np.maximum(z1, 0)
    z2 = h1 @ P["W2"].T + P["b2"]
    h2 = np.maximum(z2, 0)
    o = h2 @ P["W3"].T + P["b3"]
    return o, (x, h1, h2)


def _log_softmax(z):
    m = z.max(axis=1, keepdims=True)
    return z - m - np.log(np.exp(z - m).sum(axis=1, keepdims=True))


def loss(p, x, y_mode, y_chunk):
    o, _ = forward(p, x)
    n = len(x)
    lm = _log_softmax(o[:, :MODES])
    lc = _log_softmax(o[:, MODES:])
    return -(lm[np.arange(n), y_mode].sum() + lc[np.arange(n), y_chunk].sum()) / n


def loss_grad(p, x, y_mode, y_chunk):
    """Analytic gradient of loss() (backpropagation)."""
    P = unpack(p)
    o, (x, h1, h2) = forward(p, x)
    n = len(x)
    dO = np.zeros_like(o)
    sm = np.exp(_log_softmax(o[:, :MODES]))
    sm[np.arange(n), y_mode] -= 1
    sc = np.exp(_log_softmax(o[:, MODES:]))
    sc[np.arange(n), y_chunk] -= 1
    dO[:, :MODES] = sm / n
    dO[:, MODES:] = sc / n
    g = {"W3": dO.T @ h2, "b3": dO.sum(0)}
    dh2 = (dO @ P["W3"]) * (h2 > 0)
    g["W2"] = dh2.T @ h1
    g["b2"] = dh2.sum(0)
    dh1 = (dh2 @ P["W2"]) * (h1 > 0)
    g["W1"] = dh1.T @ x
    g["b1"] = dh1.sum(0)
    return loss(p, x, y_mode, y_chunk), pack(g)


def numeric_grad(p, x, y_mode, y_chunk, idx, eps=1e-6):
    """Central finite differences at the parameter positions `idx`."""
    g = np.zeros(len(idx))
    for k, i in enumerate(idx):
        a = p.copy()
        b = p.copy()
        a[i] += eps
        b[i] -= eps
        g[k] = (loss(a, x, y_mode, y_chunk) - loss(b, x, y_mode, y_chunk)) / (2 * eps)
    return g


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
