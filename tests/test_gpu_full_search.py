"""GPU parity of full search (dme_params.full_search, k_pel_full + the
level-0 wavefront) through the C ABI: every MEData array bit-exact against
the oracle, whose full search is pinned against the reference's own
PixelMatcher with FullSearch set (tests/test_oracle.py).  The temporal
distances come from the picture numbers, so the window radius is
min(tdiff, 3) times the range, as pixel_match.cpp:249-250 sets it."""
import numpy as np
import pytest

import oracle_py as O
import dirac_me
from dirac_me import DME_FP_C, DME_FP_MMX, DmeError, make_params, synth

pytestmark = pytest.mark.gpu


def _gpu(ctx, p, d, nums):
    cn, rn = nums
    for n, r in zip(rn, d["refs"]):
        ctx.upload(n, r)
    ctx.upload(cn, d["cur"], d["u"], d["v"])
    H, W = d["cur"].shape
    out = dirac_me.MEData(dirac_me.geometry(p, W, H))
    ctx.set_params(p)
    ctx.motion_estimate(cn, rn, out)
    for n in list(rn) + [cn]:
        ctx.drop(n)
    return out


def _same(want, got, nrefs, tag):
    for k, v in want.arrays(nrefs).items():
        g = got.arrays(nrefs)[k]
        if v.tobytes() != g.tobytes():
            raise AssertionError("%s: %s differs in %d bytes" % (tag, k, int((v.view(np.uint8) != g.view(np.uint8)).sum())))


@pytest.mark.parametrize("seed,size,nrefs,xr,yr,nums,prec,fp", [
    (60, "cif", 1, 4, 3, (1, [0]), 2, DME_FP_MMX),
    (61, "cif", 2, 6, 6, (7, [4, 9]), 2, DME_FP_MMX),      # tdiff 3 and 2: radii 18 and 12
    (62, "cif", 2, 8, 8, (10, [3, 11]), 1, DME_FP_MMX),    # tdiff 7 (capped at 3) and 1
    (63, "cif", 1, 32, 32, (5, [4]), 0, DME_FP_C),         # the encoder's default range
    (64, "cif", 2, 0, 2, (2, [1, 3]), 2, DME_FP_MMX),      # a one-row window (x range 0)
    (65, "720p", 2, 16, 12, (20, [19, 22]), 2, DME_FP_MMX),
])
def test_full_search_matches_oracle(ctx, seed, size, nrefs, xr, yr, nums, prec, fp):
    d = synth.sized(seed, size, nrefs)
    p = make_params(*d["block"], precision=prec, lam=10.0, fp=fp, full_search=True, x_range=xr, y_range=yr)
    print("case", seed, size, nrefs, xr, yr, nums, prec, fp, flush=True)
    want, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"], pic_nums=nums)
    _same(want, _gpu(ctx, p, d, nums), nrefs, "full search %s" % size)


def test_full_search_and_hierarchical_in_one_batch(ctx):
    jobs, want = [], []
    for j, (seed, full) in enumerate(((66, True), (67, False), (68, True))):
        d = synth.sized(seed, "cif", 2)
        p = make_params(*d["block"], precision=2, depth=d["depth"], lam=10.0, full_search=full, x_range=7, y_range=5)
        cn, rn = 100 * j + 5, [100 * j + 3, 100 * j + 6]
        for n, r in zip(rn, d["refs"]):
            ctx.upload(n, r)
        ctx.upload(cn, d["cur"], d["u"], d["v"])
        jobs.append((cn, rn, p))
        w, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"], pic_nums=(cn, rn))
        want.append(w)
    outs = [dirac_me.MEData(dirac_me.geometry(p, 352, 288)) for _, _, p in jobs]
    print("mixed batch", flush=True)
    ctx.run_batch(jobs, outs)
    for k, (w, o) in enumerate(zip(want, outs)):
        _same(w, o, 2, "batch job %d" % k)
    for cn, rn, _ in jobs:
        for n in [cn] + rn:
            ctx.drop(n)


def test_full_search_rejections(ctx):
    d = synth.sized(69, "cif", 1)
    ctx.upload(0, d["refs"][0])
    ctx.upload(3, d["cur"], d["u"], d["v"])
    # a window beyond one workgroup's LDS: tdiff 3, radius 384
    ctx.set_params(make_params(*d["block"], full_search=True, x_range=128, y_range=128))
    with pytest.raises(DmeError):
        ctx.motion_estimate(3, [0], None)
    # ranges without full search
    p = make_params(*d["block"])
    p.x_range_me = 8
    with pytest.raises(DmeError):
        ctx.set_params(p)
        ctx.motion_estimate(3, [0], None)
    ctx.drop(0)
    ctx.drop(3)
