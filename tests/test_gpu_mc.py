"""GPU parity of motion compensation (dme_motion_compensate and
dme_motion_compensate_batch, k_mc / k_mc_pad) through the C ABI: bit-exact
against the reference's own outputs (the mc.npz fixtures) and against the
oracle, at CIF and at full 1080p / 2160p sizes, for host arrays, device
arrays in one batched launch, and compensation straight from the device-side
motion estimate.  Full-size runs also check that subtracting and then adding
the same prediction returns the picture."""
import os
import sys

import numpy as np
import pytest

import oracle_py as O
import dirac_me
from dirac_me import DmeError, MEData, abi, geometry, make_params, synth
from dirac_me.learned import DeviceArray

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import mc_cases as M  # noqa: E402

pytestmark = pytest.mark.gpu

SUB, ADD = abi.DME_MC_SUBTRACT, abi.DME_MC_ADD


def _upload_refs(ctx, refs, base):
    nums = []
    for i, (y, u, v) in enumerate(refs):
        ctx.upload(base + i, y, u, v)
        nums.append(base + i)
    return nums


@pytest.mark.parametrize("name", [c[0] for c in M.CASES])
def test_mc_host_arrays_match_reference_fixture(ctx, name):
    print("case", name, flush=True)
    p, refs, me, planes, meta, z = M.load_case(name)
    nums = _upload_refs(ctx, refs, 100)
    ctx.motion_compensate(p, 352, 288, me, nums, meta["direction"], planes)
    for c in range(3):
        assert M.sha(planes[c]) == meta["out_sha256"][c], c


def _device_planes(ctx, planes):
    out = []
    for a in planes:
        d = DeviceArray(ctx, a.shape, np.int16)
        d.upload(a)
        out.append(d)
    return out


def test_mc_batch_of_all_fixture_cases_on_device(ctx):
    jobs, devs, metas = [], [], []
    for i, (name, *_rest) in enumerate(M.CASES):
        p, refs, me, planes, meta, z = M.load_case(name)
        nums = _upload_refs(ctx, refs, 200 + 2 * i)
        dp = _device_planes(ctx, planes)
        devs.append(dp)
        metas.append((name, meta))
        jobs.append(dict(width=352, height=288, refs=nums, direction=meta["direction"], params=p, me=me,
                         planes=[(d.ptr, a.shape[1], a.shape[0]) for d, a in zip(dp, planes)]))
    print("batch of", len(jobs), "jobs", flush=True)
    ctx.motion_compensate_batch(jobs)
    for (name, meta), dp in zip(metas, devs):
        for c in range(3):
            assert M.sha(dp[c].download()) == meta["out_sha256"][c], (name, c)


def _sized(seed, size, nrefs):
    d = synth.sized(seed, size, nrefs)
    refs = []
    for i, r in enumerate(d["refs"]):
        ru, rv = synth.chroma(r, 100 + i)
        refs.append((r, ru, rv))
    return d, refs


@pytest.mark.parametrize("size,nrefs,seed", [("1080p", 2, 31), ("720p", 1, 32)])
def test_estimate_then_compensate_on_device(ctx, size, nrefs, seed):
    """ME then MC without the MvData leaving the GPU (mv = NULL), against the
    oracle compensating with the exported MEData; then adding the prediction
    back returns the picture."""
    d, refs = _sized(seed, size, nrefs)
    H, W = d["cur"].shape
    p = make_params(*d["block"], precision=2, depth=d["depth"], lam=10.0)
    ctx.set_params(p)
    nums = _upload_refs(ctx, refs, 300)
    ctx.upload(310, d["cur"], d["u"], d["v"])
    est = MEData(geometry(p, W, H))
    print("case", size, nrefs, seed, "estimate", flush=True)
    ctx.motion_estimate(310, nums, est)
    src = [d["cur"], d["u"], d["v"]]
    dp = _device_planes(ctx, src)
    job = dict(pic_num=310, width=W, height=H, refs=nums, direction=SUB, params=p, me=None,
               planes=[(x.ptr, a.shape[1], a.shape[0]) for x, a in zip(dp, src)])
    print("compensate", flush=True)
    ctx.motion_compensate_batch([job])
    got = [x.download() for x in dp]
    want = [a.copy() for a in src]
    O.motion_compensate("oracle", p, refs, est, nrefs, SUB, want)
    for c in range(3):
        assert np.array_equal(got[c], want[c]), c
    job["direction"] = ADD
    ctx.motion_compensate_batch([job])
    for c in range(3):
        assert np.array_equal(dp[c].download(), src[c]), c
    for n in nums + [310]:
        ctx.drop(n)


@pytest.mark.parametrize("size,prec", [("2160p", 2), ("1080p", 3), ("720p", 0)])
def test_full_size_scrambled_against_oracle(ctx, size, prec):
    """Random vectors (reads far outside the picture), modes and splits, on
    padded arrays, both directions."""
    d, refs = _sized(40, size, 2)
    H, W = d["cur"].shape
    p = make_params(*d["block"], precision=2, depth=d["depth"], lam=10.0)
    me = MEData(geometry(p, W, H))
    M.scramble(me, 40, 2)
    p.mv_precision = prec
    nums = _upload_refs(ctx, refs, 400)
    src = [d["cur"], d["u"], d["v"]]
    for direction in (SUB, ADD):
        print("case", size, prec, direction, flush=True)
        got = M.padded(src, (64, 8))
        ctx.motion_compensate(p, W, H, me, nums, direction, got)
        want = M.padded(src, (64, 8))
        O.motion_compensate("oracle", p, refs, me, 2, direction, want)
        for c in range(3):
            assert np.array_equal(got[c], want[c]), (direction, c)
    for n in nums:
        ctx.drop(n)


def test_mc_rejects_bad_arguments(ctx):
    p, refs, me, planes, meta, z = M.load_case("mc_cif_q_sub")
    nums = _upload_refs(ctx, refs, 500)
    ctx.upload(510, refs[0][0])  # no chroma
    with pytest.raises(DmeError):
        ctx.motion_compensate(p, 352, 288, me, [510], SUB, planes)
    with pytest.raises(DmeError):
        ctx.motion_compensate(p, 352, 288, me, [999], SUB, planes)
    with pytest.raises(DmeError):
        ctx.motion_compensate(p, 352, 288, me, nums, 7, planes)
    small = [planes[0][:-2], planes[1], planes[2]]
    with pytest.raises(DmeError):
        ctx.motion_compensate(p, 352, 288, me, nums, SUB, [np.ascontiguousarray(a) for a in small])
    bad = make_params(*meta["block"], precision=4)
    with pytest.raises(DmeError):
        ctx.motion_compensate(bad, 352, 288, me, nums, SUB, planes)
    dp = _device_planes(ctx, planes)
    with pytest.raises(DmeError):  # no estimate for this picture number
        ctx.motion_compensate_batch([dict(pic_num=12345, width=352, height=288, refs=nums, direction=SUB, params=p,
                                          me=None, planes=[(x.ptr, a.shape[1], a.shape[0]) for x, a in zip(dp, planes)])])
    # a reference whose chroma is not 8-bit samples minus 128: its byte
    # up-converted planes could not hold it
    y0, u0, v0 = refs[0]
    u_hot = np.ascontiguousarray(u0.copy())
    u_hot[5, 7] = 200
    ctx.upload(511, y0, u_hot, v0)
    with pytest.raises(DmeError, match="motion compensation: chroma of picture 511"):
        ctx.motion_compensate(p, 352, 288, me, [511], SUB, planes)
    # the arrays are untouched by the failed calls
    _, _, _, fresh, _, _ = M.load_case("mc_cif_q_sub")
    for c in range(3):
        assert np.array_equal(planes[c], fresh[c])
    for n in nums + [510]:
        ctx.drop(n)


@pytest.mark.parametrize("prec", [0, 1, 2, 3])
def test_reference_unit_test_ref1_plus_ref2_is_twice_bipred(ctx, prec):
    """unit_tests/motion_comp_test.cpp:378-460 (testRef1and2) through the GPU,
    and bit-exact against the oracle on the same calls."""
    refs = [M.ramp_picture(0), M.ramp_picture(50)]
    nums = _upload_refs(ctx, refs, 600)
    p = make_params(12, 12, 8, 8, precision=min(prec, 2), depth=2, lam=10.0)
    p.mv_precision = prec
    g = geometry(make_params(12, 12, 8, 8, precision=2, depth=2, lam=10.0), 352, 288)
    zero = [np.zeros((288, 352), np.int16), np.zeros((144, 176), np.int16), np.zeros((144, 176), np.int16)]
    calls = [(abi.DME_REF1_ONLY, (5, 5), (5, 5)), (abi.DME_REF2_ONLY, (7, 3), (7, 3))]
    two, two_o = [a.copy() for a in zero], [a.copy() for a in zero]
    for mode, v1, v2 in calls:
        me = M.uniform_mv(MEData(g), mode, v1, v2, 1)
        ctx.motion_compensate(p, 352, 288, me, nums, ADD, two)
        O.motion_compensate("oracle", p, refs, me, 2, ADD, two_o)
    bi, bi_o = [a.copy() for a in zero], [a.copy() for a in zero]
    me = M.uniform_mv(MEData(g), abi.DME_REF1AND2, (5, 5), (7, 3), 1)
    for _ in range(2):
        ctx.motion_compensate(p, 352, 288, me, nums, ADD, bi)
        O.motion_compensate("oracle", p, refs, me, 2, ADD, bi_o)
    for c in range(3):
        assert np.array_equal(two[c], two_o[c]) and np.array_equal(bi[c], bi_o[c]), c
        assert np.abs(two[c].astype(int) - bi[c]).max() <= 5, c
    for n in nums:
        ctx.drop(n)
