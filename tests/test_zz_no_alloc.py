"""The no-allocation entry point (dme_motion_estimate_noalloc): planes in
device memory, every work buffer in the library's static scratch array.  Its
results must equal dme_motion_estimate's on the same pictures (bit-exact on
every MEData array), on two calls in a row, the second with a smaller
picture (the scratch array keeps the first call's contents); a picture whose
buffers exceed the scratch is refused before any work is queued."""
import numpy as np
import pytest

import dirac_me
from dirac_me import abi, make_params, synth
from dirac_me.learned import DeviceArray

pytestmark = pytest.mark.gpu


def _assert_same(want, got, nrefs, tag):
    for k, v in want.arrays(nrefs).items():
        g = got.arrays(nrefs)[k]
        if v.tobytes() != g.tobytes():
            n = int((v.view(np.uint8) != g.view(np.uint8)).sum())
            raise AssertionError("%s: %s differs (%d bytes)" % (tag, k, n))


def _device(ctx, arrays):
    out = []
    for a in arrays:
        x = DeviceArray(ctx, a.shape, np.int16)
        x.upload(a)
        out.append(x)
    return out


def _case(ctx, d, p, nrefs, base, tag):
    H, W = d["cur"].shape
    g = dirac_me.geometry(p, W, H)
    nums = [base, base + 2][:nrefs]
    # the existing entry point on registered pictures
    ctx.set_params(p)
    for n, r in zip(nums, d["refs"]):
        ctx.upload(n, r)
    ctx.upload(base + 1, d["cur"], d["u"], d["v"])
    want = dirac_me.MEData(g)
    ctx.motion_estimate(base + 1, nums, want)
    # the no-allocation one on device planes
    dp = _device(ctx, (d["cur"], d["u"], d["v"]))
    dr = _device(ctx, d["refs"][:nrefs])
    got = dirac_me.MEData(g)
    print("no-allocation call:", tag, flush=True)
    ctx.motion_estimate_noalloc(p, W, H, base + 1, [x.ptr for x in dp], nums, [x.ptr for x in dr], got)
    _assert_same(want, got, nrefs, tag)
    for x in dp + dr:
        x.free()
    for n in [base + 1] + nums:
        ctx.drop(n)
    return want


def test_noalloc_matches_motion_estimate_twice_second_smaller(ctx):
    d = synth.config1(3)
    p = make_params(*d["block"], precision=2, depth=4, lam=10.0)
    _case(ctx, d, p, 2, 7000, "1080p B picture, 2 refs")
    d = synth.config0(4)
    p = make_params(*d["block"], precision=2, depth=3, lam=10.0)
    _case(ctx, d, p, 1, 7010, "CIF P picture, 1 ref")


def test_noalloc_device_outputs_and_size_limit(ctx):
    d = synth.sized(5, "720p", 2)
    p = make_params(*d["block"], precision=2, depth=d["depth"], lam=10.0)
    want = _case(ctx, d, p, 2, 7020, "720p, 2 refs")
    # outputs in device memory
    H, W = d["cur"].shape
    g = dirac_me.geometry(p, W, H)
    dv = DeviceArray(ctx, (g.ynum_blocks, g.xnum_blocks), abi.MV_DTYPE)
    dsb = DeviceArray(ctx, (g.ynum_sb, g.xnum_sb), np.float32)
    ms = abi.dme_medata()
    ms.vectors[0] = dv.ptr
    ms.sb_costs = dsb.ptr
    dp = _device(ctx, (d["cur"], d["u"], d["v"]))
    dr = _device(ctx, d["refs"])
    ctx.motion_estimate_noalloc(p, W, H, 7021, [x.ptr for x in dp], [7020, 7022], [x.ptr for x in dr], ms)
    assert dv.download().tobytes() == want.vectors[0].tobytes()
    assert dsb.download().tobytes() == want.sb_costs.tobytes()
    for x in dp + dr + [dv, dsb]:
        x.free()
    # 3840x2160 with two references does not fit: refused with the size
    big = np.zeros((2160, 3840), np.int16)
    dp = _device(ctx, (big, big[:1080, :1920], big[:1080, :1920]))
    dr = _device(ctx, (big, big))
    q = make_params(24, 24, 16, 16, precision=2, depth=5, lam=10.0)
    with pytest.raises(dirac_me.DmeError) as e:
        ctx.motion_estimate_noalloc(q, 3840, 2160, 1, [x.ptr for x in dp], [0, 2], [x.ptr for x in dr], None)
    assert e.value.code == abi.DME_ERR_INVALID and "scratch" in str(e.value)
    for x in dp + dr:
        x.free()
