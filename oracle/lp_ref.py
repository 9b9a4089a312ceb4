"""CPU restatement (numpy) of the learned-proposal network and its training
kernels in dirac-research_amd/csrc/dme_learned.hip.

Test infrastructure only: the tests compare the GPU layer kernels (forward,
data gradient, weight and bias gradients, ReLU backward, the proposal loss,
Adam) with these functions on the same inputs.  The formulas themselves are
pinned by finite differences on float64 (tests/test_learned.py): nothing here
is taken from the GPU code's results.

Layout: tensors [n, h, w, 32]; a layer is w[9, 32, 32] (tap = (dy+1)*3 +
dx+1, input channel, output channel) and bias[32].  `bf16` rounds to the
nearest bfloat16 (ties to even), as the GPU stores activations and
gradients that feed the matrix cores.
"""
import numpy as np

C = 32
LAYER = 9 * C * C + C
SCALE = 8.0  # pel per unit of an offset output
TAU = 0.8  # quantile of the predicted-cost (pinball) term


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even)."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def split(params):
    """Flat parameter vector -> [(w[9,32,32], b[32])] for the four layers."""
    out = []
    for l in range(4):
        p = params[l * LAYER:(l + 1) * LAYER]
        out.append((p[:9 * C * C].reshape(9, C, C), p[9 * C * C:]))
    return out


def _shift(x, dy, dx):
    """y[.., i, j, :] = x[.., i + dy, j + dx, :], zero outside."""
    n, h, w, c = x.shape
    y = np.zeros_like(x)
    ys, ye = max(0, -dy), min(h, h - dy)
    xs, xe = max(0, -dx), min(w, w - dx)
    if ys < ye and xs < xe:
        y[:, ys:ye, xs:xe, :] = x[:, ys + dy:ye + dy, xs + dx:xe + dx, :]
    return y


def conv(x, w, b=None, relu=False, dtype=np.float32):
    """out[p][co] = b[co] + sum_tap sum_ci x[p + tap][ci] w[tap][ci][co]."""
    x = x.astype(dtype)
    out = np.zeros(x.shape[:3] + (w.shape[2],), dtype)
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        out += _shift(x, dy, dx) @ w[tap].astype(dtype)
    if b is not None:
        out += b.astype(dtype)
    return np.maximum(out, 0) if relu else out


def conv_data_grad(g, w, dtype=np.float32):
    """d in[q][ci] = sum_tap sum_co g[q - tap][co] w[tap][ci][co]."""
    g = g.astype(dtype)
    out = np.zeros(g.shape[:3] + (w.shape[1],), dtype)
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        out += _shift(g, -dy, -dx) @ w[tap].T.astype(dtype)
    return out


def conv_weight_grad(x, g, dtype=np.float64):
    """dw[tap][ci][co] = sum_p x[p + tap][ci] g[p][co]; db[co] = sum_p g[p][co]."""
    x = x.astype(dtype)
    g = g.astype(dtype)
    dw = np.zeros((9, x.shape[3], g.shape[3]), dtype)
    gf = g.reshape(-1, g.shape[3])
    for tap in range(9):
        dy, dx = tap // 3 - 1, tap % 3 - 1
        dw[tap] = _shift(x, dy, dx).reshape(-1, x.shape[3]).T @ gf
    return dw, gf.sum(axis=0)


def wta_loss(out, lab, alpha, beta, scale, dtype=np.float32):
    """Winner-takes-all proposal loss over the four hypotheses of each
    reference (channels 8r + 2h + {0,1}) plus the pinball loss of the TAU
    quantile for the predicted cost (channel 16 + r); labels [..., 8] = per
    reference (tx, ty, log1p SAD, valid).  Returns (scaled loss, scaled
    gradient)."""
    o = out.reshape(-1, C).astype(dtype)
    t = lab.reshape(-1, 8).astype(dtype)
    d = np.zeros_like(o)
    L = 0.0
    for r in range(2):
        v = t[:, 4 * r + 3] != 0
        sx = SCALE * o[:, 8 * r + 0:8 * r + 8:2] - t[:, 4 * r:4 * r + 1]
        sy = SCALE * o[:, 8 * r + 1:8 * r + 8:2] - t[:, 4 * r + 1:4 * r + 2]
        e = np.abs(sx) + np.abs(sy)
        hb = np.argmin(e, axis=1)  # first minimum
        wgt = alpha + (np.arange(4)[None, :] == hb[:, None])
        L += float(np.sum((e[np.arange(len(e)), hb] + alpha * e.sum(axis=1))[v]))
        dc = o[:, 16 + r] - t[:, 4 * r + 2]
        L += float(np.sum((beta * np.where(dc > 0, (1 - TAU) * dc, -TAU * dc))[v]))
        gx = scale * wgt * SCALE * np.sign(sx)
        gy = scale * wgt * SCALE * np.sign(sy)
        gx[~v] = 0
        gy[~v] = 0
        d[:, 8 * r + 0:8 * r + 8:2] = gx
        d[:, 8 * r + 1:8 * r + 8:2] = gy
        d[:, 16 + r] = np.where(v, scale * beta * np.where(dc > 0, 1 - TAU, np.where(dc < 0, -TAU, 0.0)), 0)
    return L * scale, d.reshape(out.shape)


def forward(x, params, rounding=True):
    """Network outputs and the activations the GPU keeps (BF16 with
    rounding; float64 without, for finite-difference checks)."""
    q = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    dt = np.float32 if rounding else np.float64
    layers = split(params)
    acts = [q(x)]
    for l, (w, b) in enumerate(layers):
        wq = q(w)
        if l < 3:
            acts.append(q(conv(acts[-1], wq, b, relu=True, dtype=dt)))
        else:
            out = conv(acts[-1], wq, b, dtype=dt)
    return out, acts


def backward(acts, dout, params, rounding=True):
    """Parameter gradient; with rounding, the GPU's BF16 roundings: the
    gradient fed to a layer's matrix products is bf16(dL/dpre), masked by
    its ReLU."""
    q = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    dt = np.float32 if rounding else np.float64
    layers = split(params)
    grads = [None] * 4
    g = q(dout)
    for l in range(3, -1, -1):
        dw, db = conv_weight_grad(acts[l], g)
        grads[l] = np.concatenate([dw.ravel(), db]).astype(np.float64)
        if l > 0:
            da = conv_data_grad(g, q(layers[l][0]), dtype=dt)
            g = q(np.where(acts[l] > 0, da, 0).astype(dt))
    return np.concatenate(grads)


def adam(p, g, m, v, lr, b1, b2, eps, t):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
