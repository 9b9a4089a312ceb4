"""GPU parity of the batched forms at the sizes they are measured at.

- The 64-picture mixed batch (16 CIF, 24 720p, 16 1080p, 8 2160p; one or
  two references by seed; block parameters per size) in one dme_run_batch.
- A batch of 16 1080p B pictures laid out as the benchmark lays them out
  (each picture with its own two references), which runs as four staggered
  parts on the four lanes.
- Stage-wise entry points on two pictures interleaved: a stage of one
  picture must read that picture's descriptors, whatever call ran last.
- Limits that a launch could not meet are refused on the host, naming the
  limit, before anything is queued.

Every job is compared bit-exactly on every MEData array with its own
single-picture oracle result (the oracle runs in threads: its C calls
release the interpreter lock)."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle_py as O
import dirac_me
from dirac_me import make_params, synth

pytestmark = pytest.mark.gpu

WORKERS = 16


def _assert_same(want, got, nrefs, tag=""):
    for k, v in want.arrays(nrefs).items():
        g = got.arrays(nrefs)[k]
        if v.tobytes() != g.tobytes():
            n = int((v.view(np.uint8) != g.view(np.uint8)).sum())
            raise AssertionError("%s %s differs (%d bytes)" % (tag, k, n))


def _oracles(cases):
    """Oracle MEData of (params, picture set, pic_nums) cases, in threads."""
    def one(c):
        p, d, nums = c
        return O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"], pic_nums=nums)[0]
    with ThreadPoolExecutor(max_workers=WORKERS) as ex:
        return list(ex.map(one, cases))


def test_mixed_batch_of_64_pictures(ctx):
    mix = synth.config2_mix(0)
    with ThreadPoolExecutor(max_workers=WORKERS) as ex:
        sets = list(ex.map(lambda m: synth.sized(m[2], m[0], m[1]), mix))
    jobs, outs, cases, nr = [], [], [], []
    for j, ((size, nrefs, seed), d) in enumerate(zip(mix, sets)):
        p = make_params(*d["block"], precision=2, depth=d["depth"], lam=10.0)
        base = 1000 + 3 * j
        nums = [base, base + 2][:nrefs]
        for n, r in zip(nums, d["refs"]):
            ctx.upload(n, r)
        ctx.upload(base + 1, d["cur"], d["u"], d["v"])
        H, W = d["cur"].shape
        jobs.append((base + 1, nums, p))
        outs.append(dirac_me.MEData(dirac_me.geometry(p, W, H)))
        cases.append((p, d, (base + 1, nums)))
        nr.append(nrefs)
    print("mixed batch: %d pictures" % len(jobs), flush=True)
    ctx.run_batch(jobs, outs)
    wants = _oracles(cases)
    for j, (w, o) in enumerate(zip(wants, outs)):
        _assert_same(w, o, nr[j], "job %d (%s, %d refs)" % (j, mix[j][0], mix[j][1]))
    for pn, nums, _ in jobs:
        for n in [pn] + list(nums):
            ctx.drop(n)


def test_1080p_batch_on_four_lanes(ctx):
    # the benchmark's layout: frames repeat every 8 jobs, every job has its
    # own uploaded references; 16 jobs split into four parts of four
    frames = [synth.config1(100 + s) for s in range(8)]
    p = make_params(*frames[0]["block"], precision=2, depth=4, lam=10.0)
    g = dirac_me.geometry(p, 1920, 1080)
    jobs, outs = [], []
    for j in range(16):
        d = frames[j % 8]
        base = 2000 + 3 * j
        ctx.upload(base, d["refs"][0])
        ctx.upload(base + 2, d["refs"][1])
        ctx.upload(base + 1, d["cur"], d["u"], d["v"])
        jobs.append((base + 1, [base, base + 2], p))
        outs.append(dirac_me.MEData(g))
    print("1080p batch of 16 on four lanes", flush=True)
    ctx.run_batch(jobs, outs)
    wants = _oracles([(p, d, (1, [0, 2])) for d in frames])
    for j, o in enumerate(outs):
        _assert_same(wants[j % 8], o, 2, "job %d" % j)
    # a second run of the same batch (buffers reused) gives the same results
    again = [dirac_me.MEData(g) for _ in jobs]
    ctx.invalidate(-1)
    ctx.run_batch(jobs, again)
    for j, (a, o) in enumerate(zip(again, outs)):
        _assert_same(o, a, 2, "rerun job %d" % j)
    for pn, nums, _ in jobs:
        for n in [pn] + list(nums):
            ctx.drop(n)


def test_stagewise_calls_on_two_pictures_interleaved(ctx):
    # pixel_search(A), pixel_search(B), then A's lambda map, sub-pel and mode
    # decision: every stage of A must use A's descriptors, not the last
    # uploaded ones (B's)
    a, b = synth.config0(11), synth.config0(12)
    p = make_params(*a["block"], precision=2, depth=3, lam=10.0)
    g = dirac_me.geometry(p, 352, 288)
    ctx.set_params(p)
    for base, d in ((3000, a), (3010, b)):
        ctx.upload(base, d["refs"][0])
        ctx.upload(base + 1, d["cur"], d["u"], d["v"])
    ctx.pixel_search(3001, [3000])
    ctx.pixel_search(3011, [3010])
    lam_a = dirac_me.MEData(g)
    ctx.set_lambda_map(3001, None, lam_a)
    want_l, _, _ = O.run_oracle(p, a["cur"], a["refs"], a["u"], a["v"], stop=O.STOP_LAMBDA)
    assert want_l.lambda_map.tobytes() == lam_a.lambda_map.tobytes()
    ctx.subpel(3001)
    got_a = dirac_me.MEData(g)
    ctx.mode_decide(3001, got_a)
    want_a, _, _ = O.run_oracle(p, a["cur"], a["refs"], a["u"], a["v"])
    _assert_same(want_a, got_a, 1, "A")
    # and B, whose stages ran interleaved with A's, is still intact
    ctx.set_lambda_map(3011, None)
    ctx.subpel(3011)
    got_b = dirac_me.MEData(g)
    ctx.mode_decide(3011, got_b)
    want_b, _, _ = O.run_oracle(p, b["cur"], b["refs"], b["u"], b["v"])
    _assert_same(want_b, got_b, 1, "B")
    for n in (3000, 3001, 3010, 3011):
        ctx.drop(n)


def test_launch_limits_are_checked_on_the_host(ctx):
    # 2048 block rows of 8: the sub-pel wavefront's per-row state cannot fit
    # one workgroup's LDS; refused before any launch, naming the limit
    y = np.zeros((16384, 128), np.int16)
    ctx.upload(4001, y)
    ctx.upload(4000, y)
    ctx.set_params(make_params(12, 12, 8, 8, precision=2, depth=3, lam=10.0))
    with pytest.raises(dirac_me.DmeError) as e:
        ctx.motion_estimate(4001, [4000])
    assert e.value.code == dirac_me.abi.DME_ERR_INVALID
    assert "block rows" in str(e.value)
    # the context is still usable
    d = synth.config0(5)
    p = make_params(*d["block"], precision=2, depth=3, lam=10.0)
    ctx.set_params(p)
    ctx.upload(4010, d["refs"][0])
    ctx.upload(4011, d["cur"], d["u"], d["v"])
    got = dirac_me.MEData(dirac_me.geometry(p, 352, 288))
    ctx.motion_estimate(4011, [4010], got)
    want, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"])
    _assert_same(want, got, 1)
    for n in (4000, 4001, 4010, 4011):
        ctx.drop(n)
