"""The GPU estimator's MEData in the reference's .imt diagnostics format: for
the pictures of the me_cif.imt fixture (whose MEData text the reference
wrote from its own estimates) and of its 1080p record (kept as a SHA-256),
dme_motion_estimate followed by dme_imt_write_picture reproduces the
reference's text byte for byte, and dme_imt_read_picture reads it back to
the same arrays."""
import os
import sys

import numpy as np
import pytest

from dirac_me import MEData, geometry, imt

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import imt_cases as I  # noqa: E402

pytestmark = pytest.mark.gpu


def _estimate(ctx, case, base):
    d, p, (num, refs) = I.inputs(case)
    H, W = d["cur"].shape
    ctx.set_params(p)
    nums = []
    for i, r in enumerate(d["refs"]):
        ctx.upload(base + i, r)
        nums.append(base + i)
    ctx.upload(base + 9, d["cur"], d["u"], d["v"])
    g = geometry(p, W, H)
    me = MEData(g)
    ctx.motion_estimate(base + 9, nums, me)
    for n in nums + [base + 9]:
        ctx.drop(n)
    return d, g, me, num, refs


def test_gpu_estimates_write_the_reference_file(ctx):
    text, index = I.load()
    out = imt.write_sequence(I.SEQUENCE["width"], I.SEQUENCE["height"], chroma=I.SEQUENCE["chroma"],
                             frame_rate=(I.SEQUENCE["frame_rate_num"], I.SEQUENCE["frame_rate_den"]),
                             pix_asr=(I.SEQUENCE["pix_asr_num"], I.SEQUENCE["pix_asr_den"]))
    out += imt.write_intra(I.INTRA_PIC)
    for k, case in enumerate(I.CIF_CASES):
        print("case", case[0], flush=True)
        d, g, me, num, refs = _estimate(ctx, case, 700 + 10 * k)
        assert I.inputs_digest(d) == index["records"][case[0]]["inputs_sha256"]
        rec = imt.write_picture(num, g, refs, me)
        assert I.sha(rec) == index["records"][case[0]]["record_sha256"], case[0]
        info, back, _ = imt.read_picture(rec)
        assert info["refs"] == refs
        for r in range(len(refs)):
            assert np.array_equal(back.vectors[r], me.vectors[r])
        assert np.array_equal(back.mode, me.mode) and np.array_equal(back.sb_split, me.sb_split)
        out += rec
    assert out == text


def test_gpu_estimate_1080p_record(ctx):
    text, index = I.load()
    print("case", I.HD_CASE[0], flush=True)
    d, g, me, num, refs = _estimate(ctx, I.HD_CASE, 760)
    want = index["records"][I.HD_CASE[0]]
    assert I.inputs_digest(d) == want["inputs_sha256"]
    rec = imt.write_picture(num, g, refs, me)
    assert len(rec) == want["record_bytes"]
    assert I.sha(rec) == want["record_sha256"]
