"""The MV-data wire format of GPU estimates: the 1080p and CIF MEData the
GPU writes are coded by dme_mvdata_encode exactly as the oracle's
restatement codes them (oracle/mvcodec_ref.py), decode back to the arrays as
the reference's decoder rebuilds them, and a batch equals its pictures."""
import numpy as np
import pytest

import mvcodec_ref as R
import dirac_me
from dirac_me import make_params, mvdata, synth

from test_mvcodec import _canonical, _lists, _same

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["cif1", "1080p"])
def test_gpu_estimates_code_and_decode(ctx, case):
    if case == "cif1":
        d = synth.config0(2)
        p = make_params(*d["block"], precision=2, depth=d["depth"], lam=d["lam"])
    else:
        d = synth.config1(108)
        p = make_params(*d["block"], precision=2, depth=4, lam=10.0)
    H, W = d["cur"].shape
    g = dirac_me.geometry(p, W, H)
    nums = [7000, 7002][: len(d["refs"])]
    for n, r in zip(nums, d["refs"]):
        ctx.upload(n, r)
    ctx.upload(7001, d["cur"], d["u"], d["v"])
    me = dirac_me.MEData(g)
    ctx.set_params(p)
    ctx.motion_estimate(7001, nums, me)
    nrefs = len(nums)
    inf = mvdata.info(g, nrefs, 2)
    data = mvdata.encode(inf, me)
    print(case, "MV data", len(data), "bytes", flush=True)
    split, mode, vec, dc = _lists(me, nrefs)
    bp = tuple(g.luma_blen[2]) + tuple(g.luma_bsep[2])
    assert data == R.encode(split, mode, vec, dc, nrefs, bp, 2)
    _, back, used = mvdata.decode(data, g, nrefs)
    assert used == len(data)
    _same(back, _canonical(me, nrefs), nrefs)
    assert mvdata.encode(inf, back) == data  # the decoded field codes to the same bytes
    assert mvdata.encode_batch([inf] * 3, [me, back, me], threads=3) == [data] * 3
    for n in [7001] + nums:
        ctx.drop(n)
