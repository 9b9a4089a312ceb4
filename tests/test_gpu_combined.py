"""GPU parity of combined-component ME (EncoderParams::CombinedME): every
stage reads the combined picture of each picture (EncPicture::DataForME /
UpDataForME(true), enc_picture.cpp:114-128; Combine, :285-367), whose
samples lie in [-128, 184] and need the 16-bit paths; SetDC still reads
luma.

- the combined picture, its pyramid and its up-converted plane equal the
  oracle's;
- every MEData array equals the reference's fixtures (combined.npz) and the
  oracle's at 1080p, at every precision and both float builds;
- combined and luma jobs sharing pictures in one batch, on several lanes and
  in a captured graph, equal their single-picture results;
- what the path cannot do is refused on the host."""
import os
import sys

import numpy as np
import pytest

import oracle_py as O
import dirac_me
from dirac_me import DME_FP_C, DME_FP_MMX, abi, make_params, synth

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import combined_cases as CC  # noqa: E402

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _assert_same(want, got, nrefs, tag=""):
    for k, v in want.arrays(nrefs).items():
        g = got.arrays(nrefs)[k]
        if v.tobytes() != g.tobytes():
            n = int((v.view(np.uint8) != g.view(np.uint8)).sum())
            raise AssertionError("%s %s differs (%d bytes)" % (tag, k, n))


@pytest.fixture(scope="module")
def cctx():
    c = dirac_me.Context(0)
    yield c
    c.close()


def _upload(ctx, d, base):
    nums = [base, base + 2][: len(d["refs_yuv"])]
    for n, r in zip(nums, d["refs_yuv"]):
        ctx.upload(n, *r)
    ctx.upload(base + 1, d["cur"], d["u"], d["v"])
    return nums


def _run(ctx, d, p, base):
    nums = _upload(ctx, d, base)
    H, W = d["cur"].shape
    got = dirac_me.MEData(dirac_me.geometry(p, W, H))
    ctx.invalidate(-1)
    ctx.run_batch([(base + 1, nums, p)], [got])
    return got, nums


def _drop(ctx, base, nrefs):
    for n in [base + 1, base, base + 2][: nrefs + 1]:
        ctx.drop(n)


def _comb(p):
    p.combined_me = 1
    return p


def test_combined_picture_pyramid_and_up_plane(cctx):
    d = CC.pictures(61, "720p", 1, bright=True)
    p = _comb(make_params(*d["block"], precision=2, depth=4, lam=10.0))
    print("combined planes 720p", flush=True)
    _run(cctx, d, p, 6000)
    H, W = d["cur"].shape
    want = O.combine(d["cur"], d["u"], d["v"])
    assert (want > 127).mean() > 0.05
    got = cctx.get_plane(6001, 3, 0, (H, W))
    assert np.array_equal(got, want)
    lv = want
    for l in range(1, 5):
        lv = O.downconvert(O.oracle(), lv, W >> l, H >> l, "oracle_downconvert")
        assert np.array_equal(cctx.get_plane(6001, 4, l, lv.shape), lv), l
    wr = O.combine(*d["refs_yuv"][0])
    assert (wr > 127).mean() > 0.05
    assert np.array_equal(cctx.get_plane(6000, 5, 0, (2 * H, 2 * W)), O.upconvert(O.oracle(), wr, "oracle_upconvert"))
    _drop(cctx, 6000, 1)


@pytest.mark.parametrize("name", ["cc_cif2_q", "cc_cif1_h", "cc_cif2_p", "cc_cif2_e", "cc_cif2_q_c", "cc_cif2_dref",
                                  "cc_cif2_q_bright", "cc_cif2_e_bright", "cc_720p2_q"])
def test_combined_me_matches_reference_fixture(cctx, name):
    import json

    z = np.load(os.path.join(GOLD, "combined.npz"))
    meta = json.loads(str(z["meta"]))[name]
    d = CC.pictures(meta["seed"], meta["size"], meta["nrefs"], meta["bright"])
    assert CC.inputs_digest(d) == meta["inputs_sha256"]
    p = _comb(make_params(*d["block"], precision=meta["precision"], depth=meta["depth"], lam=meta["lam"],
                          fp=meta["fp"]))
    print("combined fixture", name, flush=True)
    got, nums = _run(cctx, d, p, 6100)
    for k, v in got.arrays(meta["nrefs"]).items():
        assert CC.sha(v) == meta["sha256"][k], k
        if name + "/" + k in z:
            assert np.array_equal(v.view(np.uint8), z[name + "/" + k].view(np.uint8)), k
    _drop(cctx, 6100, meta["nrefs"])


@pytest.mark.parametrize("size,nrefs,prec,fp,bright", [("1080p", 2, 2, DME_FP_MMX, True),
                                                       ("1080p", 2, 3, DME_FP_MMX, False),
                                                       ("720p", 2, 1, DME_FP_C, True),
                                                       ("720p", 1, 0, DME_FP_MMX, True)])
def test_combined_me_matches_oracle(cctx, size, nrefs, prec, fp, bright):
    d = CC.pictures(62 + prec, size, nrefs, bright)
    p = _comb(make_params(*d["block"], precision=prec, depth=d["depth"], lam=10.0, fp=fp))
    print("combined vs oracle", size, nrefs, prec, fp, bright, flush=True)
    got, nums = _run(cctx, d, p, 6200)
    want, _, _ = O.run_oracle(p, d["cur"], d["refs_yuv"], d["u"], d["v"])
    _assert_same(want, got, nrefs, "%s prec %d" % (size, prec))
    _drop(cctx, 6200, nrefs)


def test_mixed_batch_of_combined_and_luma_jobs(cctx):
    # one sequence of pictures s0..s4 (CIF) and a 720p pair: combined and
    # luma jobs read the same pictures, at every precision, on four lanes
    d = CC.pictures(70, "cif", 2, bright=True)
    e = CC.pictures(71, "cif", 1, bright=False)
    seq = [d["refs_yuv"][0], (d["cur"], d["u"], d["v"]), d["refs_yuv"][1], e["refs_yuv"][0], (e["cur"], e["u"], e["v"])]
    for k, pic in enumerate(seq):
        cctx.upload(6300 + k, *pic)
    h = CC.pictures(72, "720p", 2, bright=True)
    _upload(cctx, h, 6310)
    jobs = []
    for k, (cur, refs, prec, comb) in enumerate([(1, [0, 2], 2, 1), (2, [1, 3], 3, 1), (3, [2], 1, 0), (4, [3], 0, 1),
                                               (0, [1], 2, 0)]):
        p = make_params(*d["block"], precision=prec, depth=3, lam=10.0)
        p.combined_me = comb
        jobs.append((6300 + cur, [6300 + r for r in refs], p))
    ph = _comb(make_params(*h["block"], precision=2, depth=4, lam=10.0))
    jobs.append((6311, [6310, 6312], ph))
    geoms = [dirac_me.geometry(j[2], 352, 288) for j in jobs[:-1]] + [dirac_me.geometry(ph, 1280, 720)]
    outs = [dirac_me.MEData(g) for g in geoms]
    print("mixed combined batch", flush=True)
    cctx.invalidate(-1)
    cctx.run_batch(jobs, outs)
    for (pn, refs, p), g, o in zip(jobs, geoms, outs):
        one = dirac_me.MEData(g)
        cctx.invalidate(-1)
        cctx.run_batch([(pn, refs, p)], [one])
        _assert_same(one, o, len(refs), "batch job %d" % pn)
    # the same batch captured and replayed
    gr = cctx.batch_graph(jobs)
    try:
        for rep in range(2):
            got = [dirac_me.MEData(g) for g in geoms]
            gr.launch(got)
            for (pn, refs, _), w, o in zip(jobs, outs, got):
                _assert_same(w, o, len(refs), "replay %d job %d" % (rep, pn))
    finally:
        gr.close()
    # picture 6301's combined job against the oracle
    want, _, _ = O.run_oracle(jobs[0][2], d["cur"], d["refs_yuv"], d["u"], d["v"])
    _assert_same(want, outs[0], 2, "oracle")
    for n in list(range(6300, 6305)) + [6310, 6311, 6312]:
        cctx.drop(n)


def test_combined_stagewise_entry_points(cctx):
    d = CC.pictures(73, "cif", 2, bright=True)
    p = _comb(make_params(*d["block"], precision=2, depth=3, lam=10.0))
    g = dirac_me.geometry(p, 352, 288)
    cctx.set_params(p)
    nums = _upload(cctx, d, 6400)
    cctx.invalidate(-1)
    for stop, call in ((O.STOP_PEL, lambda o: cctx.pixel_search(6401, nums, o)),
                       (O.STOP_LAMBDA, lambda o: cctx.set_lambda_map(6401, None, o)),
                       (O.STOP_SUBPEL, lambda o: cctx.subpel(6401, o)),
                       (O.STOP_MODE, lambda o: cctx.mode_decide(6401, o))):
        got = dirac_me.MEData(g)
        call(got)
        want, _, _ = O.run_oracle(p, d["cur"], d["refs_yuv"], d["u"], d["v"], stop=stop)
        keys = ["vectors1", "pred_costs1", "vectors2", "pred_costs2"] + \
            (["lambda_map"] if stop >= O.STOP_LAMBDA else []) + (list(want.arrays(2)) if stop == O.STOP_MODE else [])
        for k in keys:
            assert want.arrays(2)[k].tobytes() == got.arrays(2)[k].tobytes(), (stop, k)
    _drop(cctx, 6400, 2)


def test_combined_me_refusals(cctx):
    d = CC.pictures(74, "cif", 1)
    p = _comb(make_params(*d["block"], precision=2, depth=3, lam=10.0))

    def refused(match):
        with pytest.raises(dirac_me.DmeError) as e:
            cctx.run_batch([(6501, [6500], p)], [dirac_me.MEData(dirac_me.geometry(p, 352, 288))])
        assert e.value.code == abi.DME_ERR_INVALID and match in str(e.value), str(e.value)

    cctx.upload(6500, d["refs"][0])  # no chroma
    cctx.upload(6501, d["cur"], d["u"], d["v"])
    refused("chroma")
    u, v = d["refs_yuv"][0][1].copy(), d["refs_yuv"][0][2].copy()
    u[3, 5] = 200  # outside 8-bit range
    cctx.upload(6500, d["refs"][0], u, v)
    refused("[-128, 127]")
    cctx.upload(6500, *d["refs_yuv"][0])
    p.full_search = 1
    p.x_range_me = p.y_range_me = 8
    refused("full search")
    p.full_search = 0
    p.x_range_me = p.y_range_me = 0
    from dirac_me import learned
    cctx.set_lp_weights(learned.load_weights())
    with pytest.raises(dirac_me.DmeError) as e:
        cctx.run_batch_learned([(6501, [6500], p)], [dirac_me.MEData(dirac_me.geometry(p, 352, 288))])
    assert "luma" in str(e.value)
    got, _ = _run(cctx, d, p, 6500)  # and it runs once everything is in place
    _drop(cctx, 6500, 1)
