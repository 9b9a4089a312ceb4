"""Pins of the CPU oracle (oracle/cpu_ref.cpp).

1. Against the golden fixtures written from the reference encoder's own code
   (tests/golden/make_golden.py): bit-exact, every MEData array.
2. Against the reference build itself, live, where oracle/_ref exists.
3. Against what the mathematics fixes, independently of the oracle's code:
   constant pictures through the filters, brute-force SADs of tiny blocks,
   exact recovery of integer translations, zero motion for identical
   pictures.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import oracle_py as O
from dirac_me import DME_FP_C, DME_FP_MMX, MV_DTYPE, make_params, synth

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
import combined_cases as CC  # noqa: E402


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _fixture(name):
    z = np.load(os.path.join(GOLD, name + ".npz"))
    return json.loads(str(z["meta"])), z


def _inputs(meta):
    d = getattr(synth, meta["config"])(meta["seed"])
    digest = _sha(np.concatenate([d["cur"].ravel()] + [r.ravel() for r in d["refs"]] + [d["u"].ravel(), d["v"].ravel()]))
    assert digest == meta["inputs_sha256"], "synthetic generator changed: regenerate the fixtures"
    return d


def _params(meta, d):
    return make_params(*d["block"], precision=meta["precision"], depth=meta["depth"], lam=meta["lam"], fp=meta["fp"])


CIF_CASES = ["cif_s0_d3_q", "cif_s1_d3_q", "cif_s2_dref_q", "cif_s0_d3_h", "cif_s0_d3_p", "cif_s3_d3_q_l10.7",
             "cif_s3_d3_q_l10.7_c"]


@pytest.mark.parametrize("name", CIF_CASES)
def test_oracle_matches_reference_fixture_cif(name):
    meta, z = _fixture(name)
    d = _inputs(meta)
    p = _params(meta, d)
    got, cap, g = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"], capture=meta["depth"] > 0)
    for k, v in got.arrays(meta["nrefs"]).items():
        assert np.array_equal(v.view(np.uint8), z[k].view(np.uint8)), k
        assert _sha(v) == meta["sha256"][k], k
    if meta["depth"] > 0:
        for r in range(meta["nrefs"]):
            for l in range(1, g.depth + 1):
                assert np.array_equal(cap.level_vectors[r][l], z["level%d_vectors%d" % (l, r + 1)]), (r, l)


@pytest.mark.parametrize("name", ["hd_s0_d4_q", "hd_s0_d4_q_l10.7_c"])
def test_oracle_matches_reference_fixture_1080p(name):
    meta, z = _fixture(name)
    d = _inputs(meta)
    got, _, _ = O.run_oracle(_params(meta, d), d["cur"], d["refs"], d["u"], d["v"])
    for k, v in got.arrays(meta["nrefs"]).items():
        assert _sha(v) == meta["sha256"][k], k


def test_oracle_planes_match_reference_fixture():
    z = np.load(os.path.join(GOLD, "planes.npz"))
    for i in range(3):
        a = z["in%d" % i]
        h, w = a.shape
        assert np.array_equal(O.downconvert(O.oracle(), a, w // 2, h // 2, "oracle_downconvert"), z["down%d" % i])
        assert np.array_equal(O.upconvert(O.oracle(), a, "oracle_upconvert"), z["up%d" % i])


def test_oracle_lambda_matches_reference_fixture():
    z = np.load(os.path.join(GOLD, "lambda.npz"))
    for i in range(2):
        for n in (1, 2):
            got = O.lambda_map(O.oracle(), "oracle_lambda_map", z["v%d_0" % i], z["v%d_1" % i], 10.7, n)
            assert np.array_equal(got.view(np.uint32), z["lam%d_n%d" % (i, n)].view(np.uint32))


# ---- live comparison with the reference build -----------------------------

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")


@needs_ref
@pytest.mark.parametrize("seed", [5, 6, 7])
@pytest.mark.parametrize("prec,depth", [(2, 3), (2, -1), (1, 2), (0, 3)])
def test_oracle_matches_reference_live_cif(seed, prec, depth):
    d = synth.config0(seed)
    p = make_params(*d["block"], precision=prec, depth=depth, lam=10.0)
    want, _, _ = O.run_ref(p, d["cur"], d["refs"], d["u"], d["v"])
    got, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"])
    for k, v in want.arrays(1).items():
        assert np.array_equal(v.view(np.uint8), got.arrays(1)[k].view(np.uint8)), k


@needs_ref
@pytest.mark.parametrize("size,nrefs,seed", [("720p", 2, 11), ("720p", 1, 12)])
def test_oracle_matches_reference_live_sizes(size, nrefs, seed):
    d = synth.sized(seed, size, nrefs)
    p = make_params(*d["block"], precision=2, depth=d["depth"], lam=10.0)
    want, _, _ = O.run_ref(p, d["cur"], d["refs"], d["u"], d["v"])
    got, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"])
    for k, v in want.arrays(nrefs).items():
        assert np.array_equal(v.view(np.uint8), got.arrays(nrefs)[k].view(np.uint8)), k


@pytest.mark.skipif(not O.ref_available("c"), reason="plain-C reference build not present")
@pytest.mark.parametrize("lam", [10.0, 7.3])
def test_oracle_c_semantics_match_plain_c_reference(lam):
    d = synth.config0(9)
    p = make_params(*d["block"], precision=2, depth=3, lam=lam, fp=DME_FP_C)
    want, _, _ = O.run_ref(p, d["cur"], d["refs"], d["u"], d["v"], variant="c")
    got, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"])
    for k, v in want.arrays(1).items():
        assert np.array_equal(v.view(np.uint8), got.arrays(1)[k].view(np.uint8)), k


@needs_ref
def test_reference_search_paths_agree_at_reference_depth():
    # the harness's explicit-depth search must equal PixelMatcher::DoSearch
    d = synth.config0(4)
    g = O.geometry(make_params(*d["block"]), 352, 288)
    a, _, _ = O.run_ref(make_params(*d["block"], depth=-1), d["cur"], d["refs"], d["u"], d["v"])
    b, _, _ = O.run_ref(make_params(*d["block"], depth=g.depth), d["cur"], d["refs"], d["u"], d["v"])
    for k, v in a.arrays(1).items():
        assert np.array_equal(v.view(np.uint8), b.arrays(1)[k].view(np.uint8)), k


# ---- closed forms and invariants -------------------------------------------

@pytest.mark.parametrize("c", [-128, -37, 0, 5, 127])
def test_filters_preserve_constants(c):
    # the half-band taps sum to 256 and the up-sampling taps to 32: a flat
    # picture stays flat at every pyramid level and half-pel position
    a = np.full((40, 56), c, np.int16)
    assert np.all(O.downconvert(O.oracle(), a, 28, 20, "oracle_downconvert") == c)
    assert np.all(O.upconvert(O.oracle(), a, "oracle_upconvert") == c)


def test_upconvert_keeps_even_samples_and_clips():
    rng = np.random.default_rng(1)
    a = rng.integers(-128, 128, size=(30, 34)).astype(np.int16)
    u = O.upconvert(O.oracle(), a, "oracle_upconvert")
    assert np.array_equal(u[::2, ::2], a)
    assert u.min() >= -128 and u.max() <= 127


def test_downconvert_of_a_linear_ramp_is_the_ramp_at_half_rate():
    # interior samples of a horizontal ramp x are reproduced exactly: the
    # symmetric filter at a half-sample phase returns the midpoint 2cx+0.5
    x = np.arange(64)
    a = np.tile((x - 32).astype(np.int16), (32, 1))
    d = O.downconvert(O.oracle(), a, 32, 16, "oracle_downconvert")
    inner = d[:, 6:26]
    want = np.floor((2 * np.arange(6, 26) + 0.5 - 32) + 0.5).astype(np.int16)
    assert np.array_equal(inner, np.tile(want, (16, 1)))


def _pel_only(cur, ref, block=(12, 12, 8, 8), depth=2):
    p = make_params(*block, precision=2, depth=depth, lam=10.0)
    out, _, g = O.run_oracle(p, cur, [ref], stop=O.STOP_PEL)
    return out, g, p


def test_identical_pictures_give_zero_vectors_and_zero_sad():
    cur = synth.texture(3, 96, 80, 0)
    out, g, _ = _pel_only(cur, cur.copy())
    assert np.all(out.vectors[0]["x"] == 0) and np.all(out.vectors[0]["y"] == 0)
    inside = out.pred_costs[0]["SAD"][out.pred_costs[0]["SAD"] < 1e7]
    assert np.all(inside == 0)


@pytest.mark.parametrize("dx,dy", [(3, -2), (-5, 4), (0, 7)])
def test_integer_translation_is_recovered_exactly(dx, dy):
    # ref(x) = cur(x - d): blocks away from the borders find mv = d with SAD 0
    big = synth.texture(11, 160, 144, 0)
    cur = big[16:16 + 112, 16:16 + 128]
    ref = big[16 - dy:16 - dy + 112, 16 - dx:16 - dx + 128]
    out, g, p = _pel_only(np.ascontiguousarray(cur), np.ascontiguousarray(ref))
    v = out.vectors[0]
    s = out.pred_costs[0]["SAD"]
    inner = (slice(3, 10), slice(3, 12))
    assert np.all(v["x"][inner] == dx) and np.all(v["y"][inner] == dy)
    assert np.all(s[inner] == 0)


def test_pel_sad_brute_force_on_tiny_blocks():
    # the pel-search SAD of the chosen vector equals a direct sum over the
    # block with clamp-to-edge reference reads (me_utils.cpp:138-282)
    rng = np.random.default_rng(5)
    cur = rng.integers(-128, 128, size=(48, 64)).astype(np.int16)
    ref = rng.integers(-128, 128, size=(48, 64)).astype(np.int16)
    out, g, p = _pel_only(cur, ref, depth=1)
    xbsep = ybsep = 8
    off = 2
    for yb in range(g.ynum_blocks):
        for xb in range(g.xnum_blocks):
            x0, y0 = xb * xbsep - off, yb * ybsep - off
            xs, ys = max(x0, 0), max(y0, 0)
            xe, ye = min(x0 + 12, 64), min(y0 + 12, 48)
            if xe <= xs or ye <= ys:
                assert out.pred_costs[0]["SAD"][yb, xb] == 1e7
                continue
            mv = out.vectors[0][yb, xb]
            yy = np.clip(np.arange(ys, ye) + mv["y"], 0, 47)
            xx = np.clip(np.arange(xs, xe) + mv["x"], 0, 63)
            direct = np.abs(cur[ys:ye, xs:xe].astype(int) - ref[np.ix_(yy, xx)].astype(int)).sum()
            assert out.pred_costs[0]["SAD"][yb, xb] == direct


def test_sad_term_count_is_reported():
    d = synth.config0(0)
    p = make_params(*d["block"], precision=2, depth=3, lam=10.0)
    out, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"])
    assert 1e6 < out.sad_terms < 1e9


def _pel_terms_static(g):
    """N_sad of the pel search when the references equal the picture: every
    list but the zero list of radius min(level + 1, 5) deduplicates away, so
    the count is the sum over levels and blocks of (2r + 1)^2 Xl Yl
    (SetBlockLimits, me_utils.cpp:58-78, clipped to the level)."""
    xbl, ybl = g.luma_blen[2]
    xbs, ybs = g.luma_bsep[2]
    total = 0
    for l in range(g.depth + 1):
        W, H = g.level_width[l], g.level_height[l]
        r = min(l + 1, 5)
        for yb in range(g.level_yblocks[l]):
            y0 = yb * ybs - (ybl - ybs) // 2
            yl = min(y0 + ybl, H) - max(y0, 0)
            for xb in range(g.level_xblocks[l]):
                x0 = xb * xbs - (xbl - xbs) // 2
                xl = min(x0 + xbl, W) - max(x0, 0)
                if xl > 0 and yl > 0:
                    total += (2 * r + 1) ** 2 * xl * yl
    return total


def test_pel_sad_term_count_closed_form():
    # a static picture pins the counting rule of N_sad (the algorithmic work
    # of the pel search) independently of the oracle's search code
    from dirac_me import geometry
    d = synth.config0(2)
    p = make_params(*d["block"], precision=2, depth=3, lam=10.0)
    out, _, _ = O.run_oracle(p, d["cur"], [d["cur"].copy()], d["u"], d["v"], stop=O.STOP_PEL)
    assert out.sad_terms == _pel_terms_static(geometry(p, 352, 288))


def test_sad_term_fixture_of_the_benchmark_pictures():
    # tests/golden/nsad_1080p.json (make_nsad.py) holds the per-stage counts
    # bench.py prices its SAD-term roofline with; recount one picture, and
    # every picture stays below the analytic upper bounds of the stages
    # (every list full, no deduplication)
    z = json.load(open(os.path.join(GOLD, "nsad_1080p.json")))
    row = z["frames"][0]
    d = synth.config1(row["seed"])
    p = make_params(*d["block"], precision=2, depth=4, lam=10.0)
    pel = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"], stop=O.STOP_PEL)[0].sad_terms
    sub = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"], stop=O.STOP_SUBPEL)[0].sad_terms
    assert (pel, sub - pel) == (row["pel"], row["subpel"])
    for r in z["frames"]:
        assert 0 < r["pel"] < 458.1e6 and 0 < r["subpel"] < 98.0e6 and 0 < r["mode"] < 65.3e6


@pytest.mark.parametrize("seeds", [(5, 6), (7, 8)])
def test_intra_block_ratio_is_intra_modes_over_blocks(seeds):
    # IntraModeAnalyse (picture_compress.cpp:300-320): INTRA blocks of the
    # final Mode array over all blocks, divided in double and kept as the
    # float of MEData::SetIntraBlockRatio.  References unrelated to the
    # picture give a third of the blocks INTRA; the count is taken from the
    # reference build's own Mode array where it exists
    d, e = synth.config0(seeds[0]), synth.config0(seeds[1])
    p = make_params(*d["block"], precision=2, depth=3, lam=10.0)
    out, _, _ = O.run_oracle(p, d["cur"], [e["cur"]], d["u"], d["v"])
    modes = O.run_ref(p, d["cur"], [e["cur"]], d["u"], d["v"])[0].mode if O.ref_available() else out.mode
    n = int((modes == 0).sum())
    assert n > 100
    assert out.intra_block_ratio[0] == np.float32(n / modes.size)


def _eighth(name):
    z = np.load(os.path.join(GOLD, "eighth.npz"))
    meta = json.loads(str(z["meta"]))[name]
    d = synth.sized(meta["seed"], meta["size"], meta["nrefs"])
    digest = _sha(np.concatenate([d["cur"].ravel()] + [r.ravel() for r in d["refs"]] + [d["u"].ravel(), d["v"].ravel()]))
    assert digest == meta["inputs_sha256"], "synthetic generator changed: regenerate the fixtures"
    p = make_params(*d["block"], precision=3, depth=meta["depth"], lam=meta["lam"], fp=meta["fp"])
    return meta, z, d, p


EIGHTH = ["e_cif1_s0", "e_cif2_s3", "e_cif2_s7_l10.7", "e_cif2_s3_c", "e_720p2_s5"]


@pytest.mark.parametrize("name", EIGHTH)
def test_oracle_eighth_pel_matches_reference_fixture(name):
    # BlockDiffEighthPel (per-pixel float adds in both builds) and
    # BiBlockEighthPel with its bounds-checked read of the second reference
    # from the first reference's start (me_utils.cpp:890-1170, 1545-1785)
    meta, z, d, p = _eighth(name)
    got, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"])
    for k, v in got.arrays(meta["nrefs"]).items():
        assert _sha(v) == meta["sha256"][k], k
        if name + "/" + k in z:
            assert np.array_equal(v.view(np.uint8), z[name + "/" + k].view(np.uint8)), k


@pytest.mark.skipif(not O.ref_available(), reason="the reference build (oracle/_ref) is absent")
@pytest.mark.parametrize("seed,size,nrefs,fp", [(11, "cif", 2, DME_FP_MMX), (12, "cif", 2, DME_FP_C),
                                                (13, "cif", 1, DME_FP_MMX), (9, "720p", 2, DME_FP_C)])
def test_oracle_eighth_pel_matches_reference_live(seed, size, nrefs, fp):
    d = synth.sized(seed, size, nrefs)
    p = make_params(*d["block"], precision=3, depth=d["depth"], lam=10.0, fp=fp)
    got, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"])
    want, _, _ = O.run_ref(p, d["cur"], d["refs"], d["u"], d["v"], variant="mmx" if fp == DME_FP_MMX else "c")
    for k, v in want.arrays(nrefs).items():
        assert v.tobytes() == got.arrays(nrefs)[k].tobytes(), k


def test_mv_dtype_layout():
    assert MV_DTYPE.itemsize == 8
    assert DME_FP_MMX == 0


# ---- full search (PixelMatcher with FullSearch, pixel_match.cpp:91-155, 249-264, 365-370) ----

def _full_fixture():
    z = np.load(os.path.join(GOLD, "full_search.npz"))
    return json.loads(str(z["meta"])), z


@pytest.mark.parametrize("name", ["full_cif_s70_r6x5", "full_cif_s71_r32"])
def test_oracle_full_search_matches_reference_fixture(name):
    index, z = _full_fixture()
    m = index[name]
    d = synth.sized(m["seed"], "cif", m["nrefs"])
    digest = _sha(np.concatenate([d["cur"].ravel()] + [r.ravel() for r in d["refs"]] + [d["u"].ravel(), d["v"].ravel()]))
    assert digest == m["inputs_sha256"], "synthetic generator changed: regenerate full_search.npz"
    p = make_params(*d["block"], precision=m["precision"], lam=10.0, full_search=True, x_range=m["x_range"],
                    y_range=m["y_range"])
    got, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"], pic_nums=(m["pic_nums"][0], m["pic_nums"][1]))
    for k, v in got.arrays(m["nrefs"]).items():
        assert v.tobytes() == z[name + "/" + k].tobytes(), k


@needs_ref
@pytest.mark.parametrize("seed,nrefs,xr,yr,nums,prec,variant", [
    (72, 2, 5, 4, (7, [4, 9]), 2, "mmx"),
    (73, 2, 8, 8, (10, [3, 11]), 1, "mmx"),
    (74, 1, 3, 7, (2, [1]), 0, "c"),
    (75, 2, 0, 3, (5, [4, 6]), 2, "mmx"),
])
def test_oracle_full_search_matches_reference_live(seed, nrefs, xr, yr, nums, prec, variant):
    d = synth.sized(seed, "cif", nrefs)
    p = make_params(*d["block"], precision=prec, lam=10.0, fp=DME_FP_C if variant == "c" else DME_FP_MMX,
                    full_search=True, x_range=xr, y_range=yr)
    want, _, _ = O.run_ref(p, d["cur"], d["refs"], d["u"], d["v"], variant=variant, pic_nums=nums)
    got, _, _ = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"], pic_nums=nums)
    for k, v in want.arrays(nrefs).items():
        assert v.tobytes() == got.arrays(nrefs)[k].tobytes(), k


@pytest.mark.parametrize("dx,dy", [(21, -13), (-30, 9)])
def test_full_search_recovers_large_translations(dx, dy):
    """A picture that is its reference moved by (dx, dy) beyond the
    hierarchical search's reach but inside the full-search window: every
    block clear of the picture edges gets exactly that vector, SAD 0."""
    ref = synth.texture(80, 352, 288, nrects=0)
    cur = np.roll(ref, (-dy, -dx), axis=(0, 1))
    p = make_params(12, 12, 8, 8, precision=2, lam=10.0, full_search=True, x_range=32, y_range=32)
    me, _, g = O.run_oracle(p, cur, [ref], stop=O.STOP_PEL)
    v, pc = me.vectors[0], me.pred_costs[0]
    for y in range(4, g.ynum_blocks - 4):
        for x in range(5, g.xnum_blocks - 5):
            assert (v[y, x]["x"], v[y, x]["y"]) == (dx, dy) and pc[y, x]["SAD"] == 0, (x, y)


# ---- combined-component ME (EncoderParams::CombinedME, enc_picture.cpp:114-128, 237-367) ----

def _combined(name):
    z = np.load(os.path.join(GOLD, "combined.npz"))
    meta = json.loads(str(z["meta"]))[name]
    d = CC.pictures(meta["seed"], meta["size"], meta["nrefs"], meta["bright"])
    assert CC.inputs_digest(d) == meta["inputs_sha256"], "synthetic generator changed: regenerate the fixtures"
    p = make_params(*d["block"], precision=meta["precision"], depth=meta["depth"], lam=meta["lam"], fp=meta["fp"])
    p.combined_me = 1
    return meta, z, d, p


COMBINED = ["cc_cif2_q", "cc_cif1_h", "cc_cif2_p", "cc_cif2_e", "cc_cif2_q_c", "cc_cif2_dref", "cc_cif2_q_bright",
            "cc_cif2_e_bright", "cc_720p2_q"]


def test_combined_picture_closed_forms():
    # ValueType(sqrt(u^2 + v^2 + (y + 128)^2) - 128), truncated toward zero:
    # a Pythagorean triple is exact, a negative non-integer rounds up, the
    # brightest, most saturated sample is 184, a grey picture is itself
    y = np.full((4, 4), -128, np.int16)
    u = np.full((2, 2), 3, np.int16)
    v = np.full((2, 2), 4, np.int16)
    assert np.all(O.combine(y, u, v) == 5 - 128)
    y[:] = -127
    u[:] = 1
    v[:] = 0
    assert np.all(O.combine(y, u, v) == -126)  # sqrt(2) - 128 = -126.59
    y[:] = 127
    u[:] = -128
    v[:] = -128
    assert np.all(O.combine(y, u, v) == 184)  # sqrt(97793) - 128 = 184.7
    g = np.arange(-128, 128, dtype=np.int16).reshape(16, 16)
    z = np.zeros((8, 8), np.int16)
    assert np.array_equal(O.combine(g, z, z), g)
    # the 2 x 2 luma group shares one chroma sample (4:2:0 branch)
    rng = np.random.default_rng(5)
    y = rng.integers(-128, 128, (6, 8)).astype(np.int16)
    u = rng.integers(-128, 128, (3, 4)).astype(np.int16)
    v = rng.integers(-128, 128, (3, 4)).astype(np.int16)
    c = O.combine(y, u, v)
    for j in range(6):
        for i in range(8):
            q = int(u[j // 2, i // 2]) ** 2 + int(v[j // 2, i // 2]) ** 2 + (int(y[j, i]) + 128) ** 2
            assert c[j, i] == int(np.trunc(np.sqrt(q) - 128.0))


@pytest.mark.skipif(not O.ref_available(), reason="the reference build (oracle/_ref) is absent")
def test_combined_picture_matches_reference_live():
    rng = np.random.default_rng(6)
    for W, H in ((64, 48), (352, 288)):
        y = rng.integers(-128, 128, (H, W)).astype(np.int16)
        u = rng.integers(-128, 128, (H // 2, W // 2)).astype(np.int16)
        v = rng.integers(-128, 128, (H // 2, W // 2)).astype(np.int16)
        assert np.array_equal(O.combine(y, u, v), O.combine(y, u, v, "ref"))


@pytest.mark.parametrize("name", COMBINED)
def test_oracle_combined_me_matches_reference_fixture(name):
    meta, z, d, p = _combined(name)
    if meta["size"] == "cif":
        assert np.array_equal(O.combine(d["cur"], d["u"], d["v"]), z[name + "/combined"])
    got, _, _ = O.run_oracle(p, d["cur"], d["refs_yuv"], d["u"], d["v"])
    for k, v in got.arrays(meta["nrefs"]).items():
        assert CC.sha(v) == meta["sha256"][k], k
        if name + "/" + k in z:
            assert np.array_equal(v.view(np.uint8), z[name + "/" + k].view(np.uint8)), k


@pytest.mark.skipif(not O.ref_available(), reason="the reference build (oracle/_ref) is absent")
@pytest.mark.parametrize("seed,nrefs,prec,fp,bright", [(51, 2, 2, DME_FP_MMX, True), (52, 1, 3, DME_FP_C, True),
                                                       (53, 2, 1, DME_FP_C, False)])
def test_oracle_combined_me_matches_reference_live(seed, nrefs, prec, fp, bright):
    d = CC.pictures(seed, "cif", nrefs, bright)
    p = make_params(*d["block"], precision=prec, depth=d["depth"], lam=10.0, fp=fp)
    p.combined_me = 1
    got, _, _ = O.run_oracle(p, d["cur"], d["refs_yuv"], d["u"], d["v"])
    want, _, _ = O.run_ref(p, d["cur"], d["refs_yuv"], d["u"], d["v"], variant="mmx" if fp == DME_FP_MMX else "c")
    for k, v in want.arrays(nrefs).items():
        assert v.tobytes() == got.arrays(nrefs)[k].tobytes(), k


def test_combined_me_needs_reference_chroma():
    d = synth.sized(54, "cif", 1)
    p = make_params(*d["block"], precision=2, depth=3, lam=10.0)
    p.combined_me = 1
    with pytest.raises(RuntimeError, match="chroma"):
        O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"])
