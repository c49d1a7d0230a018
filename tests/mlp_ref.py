"""numpy restatement of the learned output-size predictor (FP32 math with the
BF16 roundings the MFMA kernels apply), for the GPU tests only."""
import numpy as np

SIZES = [(64, 24), (64, 64), (64, 64), (33, 64)]   # (out, in) per layer


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def unpack(params):
    out, off = [], 0
    for o, i in SIZES:
        W = params[off:off + o * i].reshape(o, i)
        off += o * i
        b = params[off:off + o]
        off += o
        out.append((W, b))
    assert off == params.size
    return out


def init(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    parts = []
    for o, i in SIZES:
        a = np.sqrt(6.0 / i) * scale
        parts.append(rng.uniform(-a, a, o * i))
        parts.append(rng.uniform(-0.1, 0.1, o))
    return np.concatenate(parts).astype(np.float32)


def features(ta, tb):
    """ta, tb: [n, 4, 3] float64 -> [n, 24] float32 (before BF16 rounding)."""
    allv = np.concatenate([ta, tb], axis=1)
    lo, hi = allv.min(axis=1), allv.max(axis=1)
    s = np.where(hi > lo, 1.0 / np.where(hi > lo, hi - lo, 1.0), 0.0)
    fa = (ta - lo[:, None, :]) * s[:, None, :]
    fb = (tb - lo[:, None, :]) * s[:, None, :]
    return np.concatenate([fa.reshape(-1, 12), fb.reshape(-1, 12)], axis=1).astype(np.float32)


def forward(params, x):
    """Logits [n, 33] and the BF16 activations [x, h1, h2, h3]."""
    layers = unpack(params)
    h = bf16(x)
    acts = [h]
    for k, (W, b) in enumerate(layers):
        z = (h.astype(np.float64) @ bf16(W).astype(np.float64).T + b).astype(np.float32)
        if k < 3:
            h = bf16(np.maximum(z, 0.0))
            acts.append(h)
        else:
            return z, acts


def budgets(logits, q=0.99, maximum=81):
    z = logits.astype(np.float64)
    p = np.exp(z - z.max(axis=1, keepdims=True))
    cdf = np.cumsum(p, axis=1) / p.sum(axis=1, keepdims=True)
    c = np.argmax(cdf[:, :32] >= q, axis=1)
    hit = (cdf[:, :32] >= q).any(axis=1)
    return np.where(hit, c, maximum)


def loss_and_grads(params, x, y):
    """Mean softmax cross-entropy and its gradient (flat, like params), with
    the BF16 roundings of the training kernels (activations and dZ)."""
    layers = unpack(params)
    logits, acts = forward(params, x)
    n = x.shape[0]
    z = logits.astype(np.float64)
    m = z.max(axis=1, keepdims=True)
    p = np.exp(z - m)
    s = p.sum(axis=1, keepdims=True)
    loss = float(np.mean((m[:, 0] + np.log(s[:, 0])) - z[np.arange(n), y]))
    d = p / s
    d[np.arange(n), y] -= 1.0
    d = bf16((d / n).astype(np.float32))
    grads = [None] * 4
    for k in (3, 2, 1, 0):
        W, _ = layers[k]
        h = acts[k][:, :W.shape[1]]
        grads[k] = (d.astype(np.float64).T @ h.astype(np.float64), d.astype(np.float64).sum(axis=0))
        if k > 0:
            dh = d.astype(np.float64) @ bf16(W).astype(np.float64)
            d = bf16(np.where(acts[k] > 0, dh, 0.0).astype(np.float32))
    flat = []
    for gW, gb in grads:
        flat += [gW.reshape(-1), gb]
    return loss, np.concatenate(flat)
