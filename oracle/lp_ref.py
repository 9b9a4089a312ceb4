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


# ---- exact refinement of the proposals (k_lp_refine) ------------------------

LP_BOX = 12  # the box around 4 x the level-2 vector: offsets -12 .. 12
LP_BD = 2 * LP_BOX + 1


def _win_index(dx, dy, r):
    """Index of offset (dx, dy) in AddNewVlist's window order of radius r
    (block_match.cpp:89-126): the centre, (+i, 0) and (-i, 0) for i = 1..r,
    then for each row j = 1..r the pairs (x, +j), (x, -j) for x = -r..r."""
    if dy == 0:
        return 0 if dx == 0 else (2 * dx - 1 if dx > 0 else -2 * dx)
    return 2 * r + 1 + (abs(dy) - 1) * 2 * (2 * r + 1) + (dx + r) * 2 + (1 if dy < 0 else 0)


def _block(b0, blen, bsep, boff, W, H, x, y):
    lx, ly = x * bsep[0] - boff[0], y * bsep[1] - boff[1]
    xp, yp = max(lx, 0), max(ly, 0)
    xl, yl = blen[0] - xp + lx, blen[1] - yp + ly
    if xp + xl - 1 > W - 1:
        xl = W - xp
    if yp + yl - 1 > H - 1:
        yl = H - yp
    return xp, yp, xl, yl


def refine(cur, ref, geom, r, out4, lvec2, lsad2, blocks=None):
    """Level-0 pel vectors and SADs of the learned refinement of reference r
    for the blocks (x, y) listed (all blocks if None).

    out4[k] holds the network's 32 outputs of block k (offsets of the four
    hypotheses of reference r at 8*r + 2*h, +1; the predicted log SAD at
    16 + r); lvec2/lsad2 are the plain level-2 vectors and SADs.  Each
    hypothesis is 4 x the level-2 vector g plus its offset (8 pel per unit,
    rounded to nearest even, clamped to +-10 pel); the candidates are the
    union of the +-2 windows around them, the winner the smallest
    SAD * 1024 + raster index in the box g +- 12.  If its SAD exceeds 1.2 x
    the predicted SAD, the +-4 window around g and the +-2 window around
    zero (box raster index, or 625 + window index when it lies outside the
    box) replace it where their best SAD is smaller.  Reference samples
    outside the picture repeat its edge.  Returns {(x, y): (vx, vy, sad,
    fallback, margin)}, margin the relative distance of the fallback test
    from its threshold (float rounding of expm1 may differ there)."""
    H, W = cur.shape
    X, Y = geom.xnum_blocks, geom.ynum_blocks
    X2, Y2 = lvec2.shape[1], lvec2.shape[0]
    blen, bsep = geom.luma_blen[2], geom.luma_bsep[2]
    boff = ((blen[0] - bsep[0]) // 2, (blen[1] - bsep[1]) // 2)
    pic = cur.astype(np.int32)
    rf = ref.astype(np.int32)
    res = {}
    todo = blocks if blocks is not None else [(x, y) for y in range(Y) for x in range(X)]
    for (x, y) in todo:
        k = y * X + x
        xp, yp, xl, yl = _block(None, blen, bsep, boff, W, H, x, y)
        if xl <= 0 or yl <= 0:
            res[(x, y)] = (0, 0, 10000000, False, 1.0)
            continue
        k2 = (min(y >> 2, Y2 - 1), min(x >> 2, X2 - 1))
        gx, gy = 4 * int(lvec2[k2]["x"]), 4 * int(lvec2[k2]["y"])
        o = np.asarray(out4[k], np.float32)
        hyp = []
        for h in range(4):
            hv = []
            for c in range(2):
                t = np.float32(SCALE) * o[8 * r + 2 * h + c]
                t = min(max(t, np.float32(-256.0)), np.float32(256.0))
                hv.append(int(min(max(int(np.rint(t)), 2 - LP_BOX), LP_BOX - 2)))
            hyp.append(hv)
        pred = np.expm1(np.float32(min(o[16 + r], np.float32(20.0))), dtype=np.float32)
        P = pic[yp:yp + yl, xp:xp + xl]
        # the box with clamped coordinates, then every offset's SAD
        ys = np.clip(np.arange(yp + gy - LP_BOX, yp + gy + LP_BOX + yl), 0, H - 1)
        xs = np.clip(np.arange(xp + gx - LP_BOX, xp + gx + LP_BOX + xl), 0, W - 1)
        box = rf[np.ix_(ys, xs)]
        win = np.lib.stride_tricks.sliding_window_view(box, (yl, xl))
        sads = np.abs(win - P[None, None]).sum(axis=(2, 3))  # [py + 12, px + 12]
        best = None
        for (hx, hy) in hyp:
            for py in range(hy - 2, hy + 3):
                for px in range(hx - 2, hx + 3):
                    key = int(sads[py + LP_BOX, px + LP_BOX]) * 1024 + (py + LP_BOX) * LP_BD + px + LP_BOX
                    best = key if best is None or key < best else best
        bs = best >> 10
        idx = best & 1023
        vx, vy = gx + idx % LP_BD - LP_BOX, gy + idx // LP_BD - LP_BOX
        thr = np.float32(1.2) * pred
        fb = bool(np.float32(bs) > thr)
        margin = abs(float(bs) - float(thr)) / max(float(thr), 1e-30)
        if fb:
            zin = abs(gx) <= LP_BOX - 2 and abs(gy) <= LP_BOX - 2
            k2b = None
            for py in range(-4, 5):
                for px in range(-4, 5):
                    key = int(sads[py + LP_BOX, px + LP_BOX]) * 1024 + (py + LP_BOX) * LP_BD + px + LP_BOX
                    k2b = key if k2b is None or key < k2b else k2b
            for wy in range(-2, 3):
                for wx in range(-2, 3):
                    if zin:
                        px, py = wx - gx, wy - gy
                        key = int(sads[py + LP_BOX, px + LP_BOX]) * 1024 + (py + LP_BOX) * LP_BD + px + LP_BOX
                    else:
                        zy = np.clip(np.arange(yp + wy, yp + wy + yl), 0, H - 1)
                        zx = np.clip(np.arange(xp + wx, xp + wx + xl), 0, W - 1)
                        s = int(np.abs(rf[np.ix_(zy, zx)] - P).sum())
                        key = s * 1024 + LP_BD * LP_BD + _win_index(wx, wy, 2)
                    k2b = min(k2b, key)
            if (k2b >> 10) < bs:
                bs = k2b >> 10
                i2 = k2b & 1023
                if i2 < LP_BD * LP_BD:
                    vx, vy = gx + i2 % LP_BD - LP_BOX, gy + i2 // LP_BD - LP_BOX
                else:
                    q = i2 - LP_BD * LP_BD
                    vx, vy = next((a, b) for b in range(-2, 3) for a in range(-2, 3) if _win_index(a, b, 2) == q)
        res[(x, y)] = (vx, vy, bs, fb, margin)
    return res
