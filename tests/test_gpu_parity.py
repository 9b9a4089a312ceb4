"""GPU parity: the HIP path through the C ABI against the oracle, bit-exact
on every MEData array, at the workload sizes and at the edges the reference
handles (overhanging block grids, pictures that do not move, one and two
references, every MV precision, both float semantics)."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle_py as O
import dirac_me
from dirac_me import DME_FP_C, DME_FP_MMX, make_params, synth

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu


def _upload(ctx, d):
    nums = [0, 2][: len(d["refs"])]
    for n, r in zip(nums, d["refs"]):
        ctx.upload(n, r)
    ctx.upload(1, d["cur"], d["u"], d["v"])
    return nums


def _gpu(ctx, p, d):
    ctx.set_params(p)
    nums = _upload(ctx, d)
    H, W = d["cur"].shape
    out = dirac_me.MEData(dirac_me.geometry(p, W, H))
    ctx.motion_estimate(1, nums, out)
    return out


def _assert_same(want, got, nrefs, tag=""):
    for k, v in want.arrays(nrefs).items():
        g = got.arrays(nrefs)[k]
        if v.tobytes() != g.tobytes():
            n = int((v.view(np.uint8) != g.view(np.uint8)).sum())
            raise AssertionError("%s %s differs (%d bytes)" % (tag, k, n))


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("prec", [2, 1, 0])
def test_cif_p_picture(ctx, seed, prec):
    d = synth.config0(seed)
    p = make_params(*d["block"], precision=prec, depth=d["depth"], lam=10.0)
    print("case cif seed", seed, "prec", prec, flush=True)
    want, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"])
    _assert_same(want, _gpu(ctx, p, d), 1)
    assert ctx.stats().host_lambda_jobs == 0  # lambda maps made on the device


@pytest.mark.parametrize("lam,fp", [(10.7, DME_FP_MMX), (10.7, DME_FP_C), (3.3, DME_FP_C)])
def test_cif_float_semantics(ctx, lam, fp):
    d = synth.config0(3)
    p = make_params(*d["block"], precision=2, depth=3, lam=lam, fp=fp)
    want, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"])
    _assert_same(want, _gpu(ctx, p, d), 1)


def test_cif_reference_depth(ctx):
    d = synth.config0(2)
    p = make_params(*d["block"], precision=2, depth=-1, lam=10.0)
    want, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"])
    _assert_same(want, _gpu(ctx, p, d), 1)


@pytest.mark.parametrize("lam,fp", [(10.0, DME_FP_MMX), (10.7, DME_FP_C)])
def test_1080p_b_picture(ctx, lam, fp):
    d = synth.config1(0)
    p = make_params(*d["block"], precision=2, depth=4, lam=lam, fp=fp)
    print("case 1080p lam", lam, "fp", fp, flush=True)
    want, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"])
    _assert_same(want, _gpu(ctx, p, d), 2)


@pytest.mark.parametrize("size,nrefs,seed", [("720p", 2, 21), ("720p", 1, 22), ("cif", 2, 23)])
def test_other_sizes(ctx, size, nrefs, seed):
    d = synth.sized(seed, size, nrefs)
    p = make_params(*d["block"], precision=2, depth=d["depth"], lam=10.0)
    want, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"])
    _assert_same(want, _gpu(ctx, p, d), nrefs)


def test_large_blocks_720p(ctx):
    # 24/16 blocks: 40-sample sub-superblocks and 72-sample superblocks in the
    # mode tables
    d = synth.sized(71, "720p", 2)
    p = make_params(24, 24, 16, 16, precision=2, depth=3, lam=10.0)
    want, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"])
    _assert_same(want, _gpu(ctx, p, d), 2)


def test_2160p_b_picture(ctx):
    # 3840x2160, 24/16 blocks, 5 levels, pans up to 128 pel
    d = synth.config4(0)
    p = make_params(*d["block"], precision=2, depth=d["depth"], lam=10.0)
    print("case 2160p", flush=True)
    want, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"])
    _assert_same(want, _gpu(ctx, p, d), 2)


def test_overhanging_grid_and_odd_blocks(ctx):
    # 360x240 with 10x14 blocks: rationalised to 12/16 lengths, block grid
    # overhangs on both axes, blocks entirely outside the picture
    d = synth.picture_set(31, 360, 240, 2, 6.0, 3, 10.0)
    p = make_params(10, 14, 7, 9, precision=2, depth=2, lam=10.0)
    want, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"])
    _assert_same(want, _gpu(ctx, p, d), 2)


def test_static_picture(ctx):
    cur = synth.texture(40, 352, 288, 3)
    u, v = synth.chroma(cur, 1)
    d = dict(cur=cur, u=u, v=v, refs=[cur.copy(), cur.copy()])
    p = make_params(12, 12, 8, 8, precision=2, depth=3, lam=10.0)
    want, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"])
    got = _gpu(ctx, p, d)
    _assert_same(want, got, 2)
    # no vector varies: the transition threshold is exactly 0, too close to an
    # integer for the device test, so the host made the lambda maps
    assert ctx.stats().host_lambda_jobs == 1


def test_intra_block_ratio(ctx):
    # IntraModeAnalyse on the device (k_intra_count): references unrelated to
    # the picture give many INTRA blocks; the ratio equals the oracle's and
    # the count of the device's own Mode array, for the single-picture call
    # and for a batch (the packed export)
    cases = [(synth.config0(5), synth.config0(6)), (synth.config0(7), synth.config0(8))]
    p = make_params(12, 12, 8, 8, precision=2, depth=3, lam=10.0)
    g = dirac_me.geometry(p, 352, 288)
    wants = []
    for d, e in cases:
        dd = dict(cur=d["cur"], u=d["u"], v=d["v"], refs=[e["cur"]])
        want, _, _ = O.run_oracle(p, dd["cur"], dd["refs"], dd["u"], dd["v"])
        got = _gpu(ctx, p, dd)
        _assert_same(want, got, 1)
        assert (got.mode == 0).sum() > 100
        assert got.intra_block_ratio[0] == want.intra_block_ratio[0] == np.float32((got.mode == 0).sum() / got.mode.size)
        wants.append(want)
    jobs, outs = [], []
    for j, (d, e) in enumerate(cases):
        ctx.upload(6000 + 2 * j, e["cur"])
        ctx.upload(6001 + 2 * j, d["cur"], d["u"], d["v"])
        jobs.append((6001 + 2 * j, [6000 + 2 * j], p))
        outs.append(dirac_me.MEData(g))
    ctx.run_batch(jobs, outs)
    for w, o in zip(wants, outs):
        _assert_same(w, o, 1)
        assert o.intra_block_ratio[0] == w.intra_block_ratio[0]
    for j in range(2):
        ctx.drop(6000 + 2 * j)
        ctx.drop(6001 + 2 * j)


EIGHTH = ["e_cif1_s0", "e_cif2_s3", "e_cif2_s7_l10.7", "e_cif2_s3_c", "e_720p2_s5"]


@pytest.mark.parametrize("name", EIGHTH)
def test_eighth_pel_matches_reference_fixture(ctx, name):
    # eighth-pel estimation (BlockDiffEighthPel: per-pixel float adds in both
    # builds; BiBlockEighthPel with its bounds-checked read of the second
    # reference from the first one's start) against the reference's own
    # output (tests/golden/eighth.npz)
    z = np.load(os.path.join(GOLD, "eighth.npz"))
    meta = json.loads(str(z["meta"]))[name]
    d = synth.sized(meta["seed"], meta["size"], meta["nrefs"])
    p = make_params(*d["block"], precision=3, depth=meta["depth"], lam=meta["lam"], fp=meta["fp"])
    print("eighth-pel case", name, flush=True)
    got = _gpu(ctx, p, d)
    for k, v in got.arrays(meta["nrefs"]).items():
        assert hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest() == meta["sha256"][k], k


def test_eighth_pel_1080p_and_mixed_batch(ctx):
    # a 1080p B picture at eighth pel against the oracle, and a batch that
    # mixes eighth- and quarter-pel pictures (one table launch serves both)
    d = synth.config1(106)
    p3 = make_params(*d["block"], precision=3, depth=4, lam=10.0)
    want, _, _ = O.run_oracle(p3, d["cur"], d["refs"], d["u"], d["v"])
    _assert_same(want, _gpu(ctx, p3, d), 2, "1080p eighth")
    cases = [(synth.sized(21, "cif", 2), 3), (synth.sized(22, "cif", 2), 2), (synth.sized(23, "cif", 1), 3)]
    jobs, outs, wants = [], [], []
    for j, (e, prec) in enumerate(cases):
        pe = make_params(*e["block"], precision=prec, depth=e["depth"], lam=10.0)
        base = 7000 + 3 * j
        nums = [base, base + 2][: len(e["refs"])]
        for n, r in zip(nums, e["refs"]):
            ctx.upload(n, r)
        ctx.upload(base + 1, e["cur"], e["u"], e["v"])
        jobs.append((base + 1, nums, pe))
        outs.append(dirac_me.MEData(dirac_me.geometry(pe, 352, 288)))
        wants.append(O.run_oracle(pe, e["cur"], e["refs"], e["u"], e["v"])[0])
    ctx.run_batch(jobs, outs)
    for j, (w, o) in enumerate(zip(wants, outs)):
        _assert_same(w, o, len(cases[j][0]["refs"]), "batch job %d" % j)
    for pn, nums, _ in jobs:
        for n in [pn] + nums:
            ctx.drop(n)


def test_stagewise_entry_points(ctx):
    d = synth.config0(7)
    p = make_params(*d["block"], precision=2, depth=3, lam=10.0)
    g = dirac_me.geometry(p, 352, 288)
    ctx.set_params(p)
    nums = _upload(ctx, d)
    for stop, call in ((O.STOP_PEL, lambda o: ctx.pixel_search(1, nums, o)),
                       (O.STOP_LAMBDA, lambda o: ctx.set_lambda_map(1, None, o)),
                       (O.STOP_SUBPEL, lambda o: ctx.subpel(1, o)),
                       (O.STOP_MODE, lambda o: ctx.mode_decide(1, o))):
        got = dirac_me.MEData(g)
        call(got)
        want, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"], stop=stop)
        keys = ["vectors1", "pred_costs1"] + (["lambda_map"] if stop >= O.STOP_LAMBDA else []) + \
            (list(want.arrays(1)) if stop == O.STOP_MODE else [])
        for k in keys:
            assert want.arrays(1)[k].tobytes() == got.arrays(1)[k].tobytes(), (stop, k)


def test_lambda_override(ctx):
    d = synth.config0(8)
    p = make_params(*d["block"], precision=2, depth=3, lam=10.0)
    g = dirac_me.geometry(p, 352, 288)
    lam = np.random.default_rng(3).choice(np.float32([0.0, 1.7, 10.0, 3.25]), size=(g.ynum_blocks, g.xnum_blocks))
    ctx.set_params(p)
    nums = _upload(ctx, d)
    got = dirac_me.MEData(g)
    ctx.pixel_search(1, nums)
    ctx.set_lambda_map(1, lam)
    ctx.subpel(1)
    ctx.mode_decide(1, got)
    want, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"], lambda_override=lam)
    _assert_same(want, got, 1)


def test_pyramid_and_upconverted_planes(ctx):
    d = synth.config1(1)
    p = make_params(*d["block"], precision=2, depth=4, lam=10.0)
    _gpu(ctx, p, d)
    g = dirac_me.geometry(p, 1920, 1080)
    _, cap, _ = O.run_oracle(p, d["cur"], d["refs"], stop=O.STOP_SUBPEL, capture=True)
    for l in range(1, 5):
        got = ctx.get_plane(1, 0, l, (g.level_height[l], g.level_width[l]))
        assert np.array_equal(got, cap.pyramid[l]), l
    assert np.array_equal(ctx.get_plane(0, 1, 0, (2160, 3840)), cap.ref_up[0])
    assert np.array_equal(ctx.get_plane(2, 1, 0, (2160, 3840)), cap.ref_up[1])


def test_batch_of_mixed_pictures(ctx):
    # one batched call over pictures of different sizes, reference counts and
    # block parameters; each must equal its own single-picture oracle result
    cases = [("cif", 1, 51), ("720p", 2, 52), ("cif", 2, 53), ("720p", 1, 54)]
    jobs, outs, wants = [], [], []
    for i, (size, nrefs, seed) in enumerate(cases):
        d = synth.sized(seed, size, nrefs)
        p = make_params(*d["block"], precision=2, depth=d["depth"], lam=10.0)
        base = 10 * (i + 1)
        nums = [base, base + 2][:nrefs]
        for n, r in zip(nums, d["refs"]):
            ctx.upload(n, r)
        ctx.upload(base + 1, d["cur"], d["u"], d["v"])
        H, W = d["cur"].shape
        jobs.append((base + 1, nums, p))
        outs.append(dirac_me.MEData(dirac_me.geometry(p, W, H)))
        wants.append((O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"])[0], nrefs))
    ctx.run_batch(jobs, outs)
    for (w, nrefs), o in zip(wants, outs):
        _assert_same(w, o, nrefs, "batch")


def test_batch_with_pictures_shared_between_jobs(ctx):
    # a queue of six pictures where pictures are both estimated and referenced,
    # and references are shared between jobs (so both lanes of the batch read
    # some of the same planes); each job must equal its own oracle result
    a, b = synth.sized(61, "cif", 2), synth.sized(62, "cif", 2)
    frames = [a["refs"][0], a["cur"], a["refs"][1], b["refs"][0], b["cur"], b["refs"][1]]
    chroma = [synth.chroma(f, 70 + k) for k, f in enumerate(frames)]
    for n, (f, (u, v)) in enumerate(zip(frames, chroma)):
        ctx.upload(300 + n, f, u, v)
    p = make_params(*a["block"], precision=2, depth=a["depth"], lam=10.0)
    plan = [(1, [0, 2]), (2, [0, 4]), (3, [2, 4]), (4, [3, 5]), (5, [4]), (0, [1, 5])]
    jobs, outs, wants = [], [], []
    for cur, refs in plan:
        jobs.append((300 + cur, [300 + r for r in refs], p))
        outs.append(dirac_me.MEData(dirac_me.geometry(p, 352, 288)))
        u, v = chroma[cur]
        wants.append(O.run_oracle(p, frames[cur], [frames[r] for r in refs], u, v,
                                  pic_nums=(300 + cur, [300 + r for r in refs]))[0])
    print("shared-picture batch", flush=True)
    ctx.run_batch(jobs, outs)
    for k, ((cur, refs), w, o) in enumerate(zip(plan, wants, outs)):
        _assert_same(w, o, len(refs), "job %d" % k)
    for n in range(6):
        ctx.drop(300 + n)


def test_errors_are_reported(ctx):
    p = make_params(12, 12, 8, 8, precision=2, depth=3)
    ctx.set_params(p)
    with pytest.raises(dirac_me.DmeError):
        ctx.motion_estimate(999, [0])  # unregistered picture
    with pytest.raises(dirac_me.DmeError):
        ctx.subpel(12345)  # no pel search done


@pytest.mark.gpu
def test_luma_outside_8bit_range_is_rejected(ctx):
    # samples are 8-bit values minus 128 (pic_io.cpp:396-401); the packed byte
    # SADs depend on it, so registration checks the range
    y = np.zeros((48, 64), np.int16)
    ctx.upload(500, y)  # in range: accepted
    y[7, 9] = 128
    with pytest.raises(dirac_me.DmeError):
        ctx.upload(501, y)
    y[7, 9] = -129
    with pytest.raises(dirac_me.DmeError):
        ctx.upload(502, y)
    ctx.drop(500)


def test_large_motion_leaves_byte_margin(ctx):
    # pans of about 40 pel: the search follows them out of the picture (level-0
    # vectors beyond 100 pel), so reference regions leave the replicated
    # margin of the byte planes and are copied with clamped coordinates
    cur = synth.texture(61, 352, 288, 3)
    u, v = synth.chroma(cur, 2)
    refs = [synth.warp(cur, 1.0, (44.0, -38.0), (), 1.0, 5), synth.warp(cur, 1.0, (-41.0, 36.0), (), 1.0, 6)]
    d = dict(cur=cur, u=u, v=v, refs=refs)
    p = make_params(12, 12, 8, 8, precision=2, depth=4, lam=10.0)
    want, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"])
    far = max(int(np.abs(want.arrays(2)[k][f]).max()) for k in ("vectors1", "vectors2") for f in ("x", "y"))
    assert far > 4 * 40  # quarter pel
    _assert_same(want, _gpu(ctx, p, d), 2)
