"""Training and evaluation of the learned-proposal network.

Every forward, backward and update step runs in the library's own GPU
kernels (include/dme_learned.h), on device memory the library allocates.
The training pairs are made by the plain GPU search: for each synthetic
picture it runs dme_pixel_search and dme_lp_training_pair reads out the
network inputs and the level-0 vectors as labels.

    python -m dirac_me.learned --pictures 2000 --out proposal_weights.npy
trains on seeded 1080p B pictures (2 references, the 1080p workload's
motion) and writes the weights; `evaluate` measures the Lagrangian cost
gap of learned proposals against the plain search on held-out seeds.
"""
import argparse
import ctypes as C
import os
import sys
import time

import numpy as np

from . import Context, MEData, abi, geometry, make_params, synth
from .abi import check, lib

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "proposal_weights.npy")
CH = abi.DME_LP_CHANNELS
LAYER = abi.DME_LP_LAYER_PARAMS
NPARAM = abi.DME_LP_PARAMS
FRAG = abi.DME_LP_FRAG_ELEMS


class DeviceArray:
    """A block of device memory of the context's device (dme_lp_alloc)."""

    def __init__(self, ctx, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(x) for x in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().dme_lp_alloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p.value

    def at(self, byte_offset=0):
        return C.c_void_p(self.ptr + byte_offset)

    def slice_ptr(self, i):
        """Pointer to element i of the first axis."""
        return self.at(i * (self.nbytes // self.shape[0]))

    def upload(self, a, byte_offset=0):
        a = np.ascontiguousarray(a, self.dtype)
        check(lib().dme_lp_copy(self.ctx.h, self.at(byte_offset), a.ctypes.data_as(C.c_void_p), a.nbytes, 1))

    def download(self):
        out = np.empty(self.shape, self.dtype)
        check(lib().dme_lp_copy(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.at(), self.nbytes, 2))
        return out

    def zero(self):
        check(lib().dme_lp_zero(self.ctx.h, self.at(), self.nbytes))

    def free(self):
        if self.ptr:
            lib().dme_lp_free(self.ctx.h, C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def init_params(seed=0):
    """He-normal 3x3 weights over 9 x 32 inputs, zero biases; the last layer
    starts small so that the first proposals sit on the level-2 guide."""
    rng = np.random.default_rng(seed)
    out = []
    for l in range(4):
        w = rng.standard_normal((9, CH, CH)).astype(np.float32) * np.sqrt(2.0 / (9 * CH))
        if l == 0:
            w[:, 8:, :] = 0.0  # unused input channels
        if l == 3:
            w *= 0.1
        out.append(np.concatenate([w.ravel(), np.zeros(CH, np.float32)]))
    return np.concatenate(out).astype(np.float32)


def load_weights(path=WEIGHTS):
    p = np.load(path, allow_pickle=False)
    if p.dtype != np.float32 or p.shape != (NPARAM,):
        raise ValueError("%s: expected %d float32 parameters" % (path, NPARAM))
    return p


class Network:
    """Parameters, gradients and Adam moments in device memory."""

    def __init__(self, ctx, params):
        self.ctx = ctx
        self.p = DeviceArray(ctx, NPARAM, np.float32)
        self.p.upload(np.asarray(params, np.float32))
        self.g = DeviceArray(ctx, NPARAM, np.float32)
        self.m = DeviceArray(ctx, NPARAM, np.float32)
        self.v = DeviceArray(ctx, NPARAM, np.float32)
        self.frag = DeviceArray(ctx, 4 * 2 * FRAG, np.uint16)
        self.loss = DeviceArray(ctx, 1, np.float32)
        for a in (self.g, self.m, self.v, self.loss):
            a.zero()
        self.t = 0
        self._bufs = {}
        self.prepare()

    def _layer(self, l):
        w = self.p.ptr + 4 * l * LAYER
        return C.c_void_p(w), C.c_void_p(w + 4 * 9 * CH * CH)

    def _frag(self, l, bwd):
        return self.frag.at(2 * (l * 2 + (1 if bwd else 0)) * FRAG)

    def prepare(self):
        for l in range(4):
            w, _ = self._layer(l)
            check(lib().dme_lp_prepare(self.ctx.h, w, self._frag(l, False), self._frag(l, True)))

    def _buf(self, npos):
        if self._bufs.get("npos", 0) < npos:
            self._bufs = dict(
                npos=npos,
                act=[DeviceArray(self.ctx, npos * CH, np.uint16) for _ in range(3)],
                out=DeviceArray(self.ctx, npos * CH, np.float32),
                dout=DeviceArray(self.ctx, npos * CH, np.float32),
                gbf=DeviceArray(self.ctx, npos * CH, np.uint16),
                da=DeviceArray(self.ctx, npos * CH, np.float32),
            )
        return self._bufs

    def forward(self, feat_ptr, n, h, w):
        """Network outputs (f32) of n grids of h x w positions; the BF16
        activations stay in the buffers for a backward pass."""
        b = self._buf(n * h * w)
        acts = [feat_ptr] + [a.at() for a in b["act"]]
        for l in range(4):
            wp, bp = self._layer(l)
            if l < 3:
                check(lib().dme_lp_conv(self.ctx.h, acts[l], n, h, w, self._frag(l, False), bp, 1, acts[l + 1], None))
            else:
                check(lib().dme_lp_conv(self.ctx.h, acts[l], n, h, w, self._frag(l, False), bp, 0, None,
                                        b["out"].at()))
        return b["out"], acts

    def backward(self, acts, n, h, w):
        """Parameter gradient into self.g from the loss gradient in dout."""
        b = self._bufs
        npos = n * h * w
        self.g.zero()
        check(lib().dme_lp_relu_backward(self.ctx.h, None, b["dout"].at(), npos * CH, b["gbf"].at()))
        for l in range(3, -1, -1):
            gw = self.g.at(4 * l * LAYER)
            gb = self.g.at(4 * (l * LAYER + 9 * CH * CH))
            check(lib().dme_lp_wgrad(self.ctx.h, acts[l], b["gbf"].at(), n, h, w, gw, gb))
            if l > 0:
                check(lib().dme_lp_conv(self.ctx.h, b["gbf"].at(), n, h, w, self._frag(l, True), None, 0, None,
                                        b["da"].at()))
                check(lib().dme_lp_relu_backward(self.ctx.h, acts[l], b["da"].at(), npos * CH, b["gbf"].at()))

    def loss_and_grad(self, feat_ptr, lab_ptr, n, h, w, alpha=0.02, beta=1.0):
        npos = n * h * w
        out, acts = self.forward(feat_ptr, n, h, w)
        self.loss.zero()
        check(lib().dme_lp_wta_loss(self.ctx.h, out.at(), lab_ptr, npos, alpha, beta, 1.0 / npos,
                                    self._bufs["dout"].at(), self.loss.at()))
        self.backward(acts, n, h, w)

    def loss_value(self):
        return float(self.loss.download()[0])

    def step(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.t += 1
        check(lib().dme_lp_adam(self.ctx.h, self.p.at(), self.g.at(), self.m.at(), self.v.at(), NPARAM, lr, b1, b2,
                                eps, self.t))
        self.prepare()

    def params(self):
        return self.p.download().copy()


def _picture_seed(i, held_out):
    return (1_000_003 if held_out else 7) + 7919 * i


def _make_set(args):
    i, held_out, size = args
    if size == "1080p":
        return synth.config1(_picture_seed(i, held_out))
    return synth.sized(_picture_seed(i, held_out), size, 2)


def pictures(count, held_out=False, size="1080p", workers=8):
    """Seeded picture sets, generated on host threads a few ahead, in order."""
    from collections import deque
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(max_workers=workers) as ex:
        q = deque()
        nxt = 0
        while nxt < count or q:
            while nxt < count and len(q) < 2 * workers:
                q.append(ex.submit(_make_set, (nxt, held_out, size)))
                nxt += 1
            yield q.popleft().result()


def build_dataset(ctx, count, size="1080p", log=None):
    """Training pairs of `count` pictures: features [count, h, w, 32] (BF16)
    and labels [count, h, w, 8], in device memory."""
    feats = labs = None
    t0 = time.time()
    for i, d in enumerate(pictures(count, False, size)):
        if feats is None:
            prm = make_params(*d["block"], precision=d["precision"], depth=d["depth"], lam=d["lam"])
            ctx.set_params(prm)
            g = geometry(prm, d["cur"].shape[1], d["cur"].shape[0])
            h, w = g.ynum_blocks, g.xnum_blocks
            feats = DeviceArray(ctx, (count, h, w, CH), np.uint16)
            labs = DeviceArray(ctx, (count, h, w, 8), np.float32)
        ctx.upload(0, d["refs"][0])
        ctx.upload(2, d["refs"][1])
        ctx.upload(1, d["cur"], d["u"], d["v"])
        ctx.pixel_search(1, [0, 2])
        check(lib().dme_lp_training_pair(ctx.h, 1, feats.slice_ptr(i), labs.slice_ptr(i)))
        ctx.invalidate(-1)
        if log and (i + 1) % 100 == 0:
            log("dataset: %d/%d pictures, %.1f s" % (i + 1, count, time.time() - t0))
    ctx.synchronize()
    return feats, labs


def train(ctx, feats, labs, epochs=12, batch=8, lr=2e-3, seed=0, params=None, log=None):
    net = Network(ctx, init_params(seed) if params is None else params)
    n, h, w = feats.shape[:3]
    per_f, per_l = feats.nbytes // n, labs.nbytes // n
    bf = DeviceArray(ctx, (batch, h, w, CH), np.uint16)
    bl = DeviceArray(ctx, (batch, h, w, 8), np.float32)
    rng = np.random.default_rng(seed)
    steps = 0
    for ep in range(epochs):
        order = rng.permutation(n)
        tot = 0.0
        nb = 0
        lr_ep = lr * (0.5 ** max(0, ep - epochs // 2))
        for s0 in range(0, n - batch + 1, batch):
            for k, i in enumerate(order[s0:s0 + batch]):
                check(lib().dme_lp_copy(ctx.h, bf.at(k * per_f), feats.slice_ptr(int(i)), per_f, 3))
                check(lib().dme_lp_copy(ctx.h, bl.at(k * per_l), labs.slice_ptr(int(i)), per_l, 3))
            net.loss_and_grad(bf.at(), bl.at(), batch, h, w)
            net.step(lr_ep)
            tot += net.loss_value()
            nb += 1
            steps += 1
        if log:
            log("epoch %d: mean loss %.4f (lr %.2e, %d steps)" % (ep, tot / max(1, nb), lr_ep, steps))
    return net.params()


def sb_cost_total(me):
    return float(np.sum(me.sb_costs.astype(np.float64)))


def evaluate(ctx, params, count=8, size="1080p", log=None):
    """Mean Lagrangian cost gap (sum of SBCosts, learned / plain - 1) and mean
    pel-cost gap over `count` held-out seeds, with the fallback rate."""
    ctx.set_lp_weights(params)
    gaps, pel_gaps = [], []
    b0, f0 = ctx.lp_stats()
    for d in pictures(count, True, size):
        prm = make_params(*d["block"], precision=d["precision"], depth=d["depth"], lam=d["lam"])
        ctx.set_params(prm)
        g = geometry(prm, d["cur"].shape[1], d["cur"].shape[0])
        ctx.upload(0, d["refs"][0])
        ctx.upload(2, d["refs"][1])
        ctx.upload(1, d["cur"], d["u"], d["v"])
        plain, learned = MEData(g), MEData(g)
        ctx.motion_estimate(1, [0, 2], plain)
        ctx.motion_estimate_learned(1, [0, 2], learned)
        gaps.append(sb_cost_total(learned) / sb_cost_total(plain) - 1.0)
        tot = lambda m: float(sum(np.sum(m.pred_costs[r]["total"].astype(np.float64)) for r in range(2)))
        pel_gaps.append(tot(learned) / tot(plain) - 1.0)
        ctx.invalidate(-1)
    b1, f1 = ctx.lp_stats()
    res = dict(gap=float(np.mean(gaps)), gaps=[float(x) for x in gaps], pel_gap=float(np.mean(pel_gaps)),
               fallback_rate=(f1 - f0) / max(1, b1 - b0))
    if log:
        log("evaluate: mean SB-cost gap %.4f, pel-cost gap %.4f, fallback %.3f" % (
            res["gap"], res["pel_gap"], res["fallback_rate"]))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pictures", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval", type=int, default=8)
    ap.add_argument("--out", default=WEIGHTS)
    args = ap.parse_args()

    def log(msg):
        print(msg, flush=True)

    ctx = Context(0)
    feats, labs = build_dataset(ctx, args.pictures, log=log)
    params = train(ctx, feats, labs, args.epochs, args.batch, args.lr, log=log)
    np.save(args.out, params.astype(np.float32), allow_pickle=False)
    log("weights: %s" % args.out)
    feats.free()
    labs.free()
    evaluate(ctx, init_params(0), args.eval, log=lambda m: log("untrained " + m))
    evaluate(ctx, params, args.eval, log=lambda m: log("trained " + m))
    ctx.close()


if __name__ == "__main__":
    sys.exit(main())
