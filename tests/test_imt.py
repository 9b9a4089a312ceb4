"""The .imt diagnostics text through the C ABI (dme_imt_write_sequence,
dme_imt_write_picture, dme_imt_read_sequence, dme_imt_read_picture), and the
oracle's restatement of its writer (oracle_py.imt_*).

Pins:
1. A known-answer record written out by hand from encmain.cpp:206-330 and
   motion.cpp:452-485, whose floats cover the "%.6g" cases (integers, -0,
   exponent forms, rounding to six digits).
2. Against the reference's own operator<<(ostream&, MEData&) and its reader
   (ProcessSequence::AddPictureEntry with operator>> for OLBParams and
   MEData), live, where oracle/_ref exists: random MEData, one and two
   references, several shapes; bit-exact arrays, byte-exact text.
3. The fixture me_cif.imt, whose MEData text the reference wrote from its own
   estimates: read back, and rewritten byte for byte from the oracle's
   estimates of the same pictures.
4. Round trips: ints exact, floats to the float nearest their six-digit
   decimal.
"""
import ctypes as C
import os
import sys
import types

import numpy as np
import pytest

import oracle_py as O
from dirac_me import COST_DTYPE, DmeError, MEData, abi, imt

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import imt_cases as I  # noqa: E402

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")


def _me(by, bx, sy, sx):
    return MEData(types.SimpleNamespace(ynum_blocks=by, xnum_blocks=bx, ynum_sb=sy, xnum_sb=sx))


def _geom(by, bx, sy, sx, ybsep=8, xbsep=8):
    g = abi.dme_geometry()
    g.ynum_blocks, g.xnum_blocks, g.ynum_sb, g.xnum_sb = by, bx, sy, sx
    g.luma_bsep[2][0], g.luma_bsep[2][1] = xbsep, ybsep
    return g


def _costs(pairs):
    a = np.zeros((len(pairs), len(pairs[0])), COST_DTYPE)
    for j, row in enumerate(pairs):
        for i, (s, m) in enumerate(row):
            a[j, i]["SAD"], a[j, i]["mvcost"], a[j, i]["total"] = s, m, 777.0
    return a


def _vectors(pairs):
    a = np.zeros((len(pairs), len(pairs[0])), abi.MV_DTYPE)
    for j, row in enumerate(pairs):
        for i, (x, y) in enumerate(row):
            a[j, i]["x"], a[j, i]["y"] = x, y
    return a


def _known():
    me = _me(2, 3, 1, 2)
    me.sb_split[...] = [[0, 2]]
    me.sb_costs[...] = [[1.5, 1234567.0]]
    me.mode[...] = [[0, 1, 3], [2, 3, 1]]
    me.intra_costs[...] = [[0.0, -0.0, 100000.0], [123456.7, 3.14159265, 1e-5]]
    me.bipred_costs[...] = _costs([[(10, 0.25), (7.125, 2), (1e6, 0.0001)],
                                   [(12345678, 5.5), (-1, 0), (0.1, 99.99999)]])
    me.dc[0][...] = [[-128, 0, 127], [5, -5, 300]]
    me.dc[1][...] = [[1, 2, 3], [4, 5, 6]]
    me.dc[2][...] = [[-1, -2, -3], [-4, -5, -6]]
    me.vectors[0][...] = _vectors([[(0, 0), (-3, 4), (12, -7)], [(100, -100), (1, 1), (-1, -1)]])
    me.pred_costs[0][...] = _costs([[(1, 2), (3, 4), (5, 6)], [(7, 8), (9, 10), (11.5, 12.25)]])
    me.vectors[1][...] = _vectors([[(2, 2), (0, -1), (-9, 0)], [(3, 3), (4, 4), (5, 5)]])
    me.pred_costs[1][...] = _costs([[(0.5, 0), (1e-7, 1e7), (2, 2)], [(4096, 0.125), (65536, 1), (999999.5, 3)]])
    return me


# WriteDiagnosticsData (encmain.cpp:206-330) for the record above, by hand.
KNOWN = "\n".join([
    "",
    "[frame:7]>mo_comp",
    "",
    "2 3 11 8 8 1 2 2 3",
    "",
    "0 2 ",
    "",
    "1.5 1.23457e+06 ",
    "",
    "0 1 3 ",
    "2 3 1 ",
    "",
    "0 -0 100000 ",
    "123457 3.14159 1e-05 ",
    "",
    "10 0.25 7.125 2 1e+06 0.0001 ",
    "1.23457e+07 5.5 -1 0 0.1 100 ",
    "",
    "-128 0 127 ",
    "5 -5 300 ",
    "",
    "1 2 3 ",
    "4 5 6 ",
    "",
    "-1 -2 -3 ",
    "-4 -5 -6 ",
    "",
    "0 0 -3 4 12 -7 ",
    "100 -100 1 1 -1 -1 ",
    "",
    "1 2 3 4 5 6 ",
    "7 8 9 10 11.5 12.25 ",
    "",
    "",
    "2 2 0 -1 -9 0 ",
    "3 3 4 4 5 5 ",
    "",
    "0.5 0 1e-07 1e+07 2 2 ",
    "4096 0.125 65536 1 1e+06 3 ",
    "",
    "",
]).encode()


def test_known_answer_record():
    me = _known()
    assert imt.write_picture(7, _geom(2, 3, 1, 2), [3, 11], me) == KNOWN
    assert O.imt_record(7, [3, 11], 8, 8, me) == KNOWN


def test_known_answer_one_reference_intra_and_sequence():
    me = _known()
    text = imt.write_picture(5, _geom(2, 3, 1, 2, ybsep=12, xbsep=16), [4], me)
    lines = text.decode().split("\n")
    assert lines[:4] == ["", "[frame:5]>mo_comp", "", "1 4 12 16 1 2 2 3"]
    # no BiPredCosts: after the intra-cost rows two empty lines, then DC(Y_COMP)
    i = lines.index("123457 3.14159 1e-05 ")
    assert lines[i + 1:i + 4] == ["", "", "-128 0 127 "]
    assert text.endswith(b"\n7 8 9 10 11.5 12.25 \n\n")
    assert text == O.imt_record(5, [4], 12, 16, me)
    assert imt.write_intra(9) == b"\n[frame:9]>intra\n" == O.imt_intra(9)
    seq = imt.write_sequence(1920, 1080, chroma=2, source_sampling=0, topfieldfirst=1, frame_rate=(50, 1),
                             pix_asr=(1, 1), picture_coding_mode=1)
    assert seq == b"2\n1920\n1080\n0\n1\n50\n1\n1\n1\n1\n"
    assert imt.read_sequence(seq) == (dict(chroma=2, width=1920, height=1080, source_sampling=0, topfieldfirst=1,
                                           frame_rate_num=50, frame_rate_den=1, pix_asr_num=1, pix_asr_den=1,
                                           picture_coding_mode=1), len(seq))


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def _same_read(got, want, nrefs):
    """Bit-exact equality of everything the text carries."""
    for k in ["sb_split", "sb_costs", "mode", "intra_costs"] + (["bipred_costs"] if nrefs > 1 else []):
        a, b = getattr(got, k), getattr(want, k)
        if a.dtype == COST_DTYPE:
            for f in ("SAD", "mvcost"):
                assert np.array_equal(_bits(a[f]), _bits(b[f])), (k, f)
        else:
            assert np.array_equal(_bits(a), _bits(b)), k
    for c in range(3):
        assert np.array_equal(got.dc[c], want.dc[c]), c
    for r in range(nrefs):
        assert np.array_equal(got.vectors[r], want.vectors[r]), r
        for f in ("SAD", "mvcost"):
            assert np.array_equal(_bits(got.pred_costs[r][f]), _bits(want.pred_costs[r][f])), (r, f)


def _random_me(seed, by, bx, sy, sx):
    rng = np.random.default_rng(seed)
    me = _me(by, bx, sy, sx)

    def floats(shape):
        # magnitudes from 1e-8 to 1e9, integers, halves, zeros of both signs
        m = rng.uniform(-8, 9, shape)
        v = (10.0 ** m) * rng.choice([-1, 1], shape)
        v = np.where(rng.random(shape) < 0.2, np.round(v), v)
        v = np.where(rng.random(shape) < 0.05, rng.choice([0.0, -0.0], shape), v)
        return v.astype(np.float32)

    me.sb_split[...] = rng.integers(0, 3, (sy, sx))
    me.sb_costs[...] = floats((sy, sx))
    me.mode[...] = rng.integers(0, 4, (by, bx))
    me.intra_costs[...] = floats((by, bx))
    for a in [me.bipred_costs] + me.pred_costs:
        for f in ("SAD", "mvcost", "total"):
            a[f] = floats((by, bx))
    for a in me.dc:
        a[...] = rng.integers(-32768, 32768, (by, bx))
    for v in me.vectors:
        v["x"] = rng.integers(-2 ** 31, 2 ** 31, (by, bx))
        v["y"] = rng.integers(-600, 600, (by, bx))
    return me


SHAPES = [(4, 4, 1, 1), (36, 44, 9, 11), (3, 17, 1, 5), (68, 120, 17, 30)]


@needs_ref
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("nrefs", [1, 2])
def test_writer_and_reader_match_the_reference(shape, nrefs):
    me = _random_me(sum(shape) * 7 + nrefs, *shape)
    by, bx, sy, sx = shape
    text = imt.write_picture(11, _geom(*shape, ybsep=12, xbsep=8), [5, 13][:nrefs], me)
    head = O.imt_record_head(11, [5, 13][:nrefs], 12, 8, (sy, sx), (by, bx))
    want_body = O.ref_imt_medata_text(me, nrefs)
    assert text == head + want_body
    assert O.imt_medata_text(me, nrefs) == want_body
    info, got, off = imt.read_picture(text)
    rinfo, want = O.ref_imt_read(text)
    assert info == rinfo and off == len(text)
    assert info == dict(pic_num=11, intra=False, refs=[5, 13][:nrefs], ybsep=12, xbsep=8, sb_shape=(sy, sx),
                        mv_shape=(by, bx))
    _same_read(got, want, nrefs)


@needs_ref
def test_reader_matches_the_reference_on_intra_and_fixture_records():
    text, index = I.load()
    seq, off = imt.read_sequence(text)
    assert seq == I.SEQUENCE
    n = 0
    while off < len(text):
        info, me, nxt = imt.read_picture(text, off)
        rinfo, rme = O.ref_imt_read(text[off:])
        assert info == rinfo
        if me is not None:
            _same_read(me, rme, len(info["refs"]))
        off, n = nxt, n + 1
    assert n == 1 + len(I.CIF_CASES)


def test_round_trip():
    for nrefs in (1, 2):
        me = _random_me(nrefs, 9, 13, 3, 4)
        text = imt.write_picture(3, _geom(9, 13, 3, 4), [1, 2][:nrefs], me)
        _, got, _ = imt.read_picture(text)
        six = np.vectorize(lambda x: np.float32(float("%.6g" % x)), otypes=[np.float32])
        for k in ("sb_costs", "intra_costs"):
            assert np.array_equal(_bits(getattr(got, k)), _bits(six(getattr(me, k)))), k
        for a, b in [(got.bipred_costs, me.bipred_costs)][:nrefs - 1] + list(zip(got.pred_costs, me.pred_costs))[
                :nrefs]:
            for f in ("SAD", "mvcost"):
                assert np.array_equal(_bits(a[f]), _bits(six(b[f]))), f
            assert not a["total"].any()  # not in the text
        assert np.array_equal(got.mode, me.mode) and np.array_equal(got.sb_split, me.sb_split)
        for r in range(nrefs):
            assert np.array_equal(got.vectors[r], me.vectors[r])
        for c in range(3):
            assert np.array_equal(got.dc[c], me.dc[c])
        # what was read writes the same text again
        assert imt.write_picture(3, _geom(9, 13, 3, 4), [1, 2][:nrefs], got) == text


def test_fixture_reads_and_is_rewritten_from_the_oracle_estimates():
    """me_cif.imt's MEData text came from the reference; the oracle's
    estimates of the same pictures, written by the product writer, give the
    same file byte for byte."""
    text, index = I.load()
    assert I.sha(text) == index["file_sha256"]
    out = imt.write_sequence(I.SEQUENCE["width"], I.SEQUENCE["height"], chroma=I.SEQUENCE["chroma"],
                             source_sampling=0, topfieldfirst=1, frame_rate=(25, 1), pix_asr=(12, 11))
    out += imt.write_intra(I.INTRA_PIC)
    for case in I.CIF_CASES:
        d, p, nums = I.inputs(case)
        assert I.inputs_digest(d) == index["records"][case[0]]["inputs_sha256"], "regenerate the .imt fixture"
        me, _, g = O.run_oracle(p, d["cur"], d["refs"], d["u"], d["v"], pic_nums=nums)
        out += imt.write_picture(nums[0], g, nums[1], me)
    assert out == text
    seq, recs = imt.read(text)
    assert seq == I.SEQUENCE and [r[0]["pic_num"] for r in recs] == [0, 1, 2]
    assert recs[0][1] is None and recs[0][0]["intra"]


def test_float_formatting_dense_against_printf():
    """Floats across the whole range (the writer's integer fast path covers
    the fixed forms, the general conversion the rest) against Python's
    correctly rounded "%.6g", which is C's."""
    rng = np.random.default_rng(5)
    bits = np.concatenate([rng.integers(0, 2 ** 31, 150000, dtype=np.uint64),          # any positive float
                           rng.integers(0x38D1B717, 0x49742400, 150000, dtype=np.uint64)])  # 1e-4 .. 1e6
    v = bits.astype(np.uint32).view(np.float32)
    v = v[np.isfinite(v)]
    v = np.concatenate([v, -v[:1000], np.float32([0.0, -0.0, 1e-4, 999999.5, 999999.4, 99999.95, 0.00012345675])])
    n = len(v)
    me = _me(1, n, 1, 1)
    me.intra_costs[0] = v
    text = imt.write_picture(0, _geom(1, n, 1, 1), [1], me).decode().split("\n")
    assert text[9] == "0 " * n  # the mode row; a blank line, then the intra costs
    got = text[11].split(" ")[:-1]
    want = ["%.6g" % float(x) for x in v]
    assert got == want


def test_record_bound_holds_for_the_widest_values():
    me = _me(3, 5, 1, 2)
    for a in (me.sb_costs, me.intra_costs):
        a[...] = np.float32(-1.23457e-05)
    for a in [me.bipred_costs] + me.pred_costs:
        a["SAD"] = a["mvcost"] = np.float32(-0.000123457)
    me.mode[...] = me.sb_split[...] = -2 ** 31
    for a in me.dc:
        a[...] = -32768
    for v in me.vectors:
        v["x"] = v["y"] = -2 ** 31
    for nrefs in (1, 2):
        p = imt.picture_info(-2 ** 31, _geom(3, 5, 1, 2, ybsep=-2 ** 31, xbsep=-2 ** 31), [-2 ** 31] * nrefs)
        n = C.c_size_t()
        assert abi.lib().dme_imt_write_picture(C.byref(p), C.byref(me.struct()), None, 0, C.byref(n)) == 0
        assert n.value <= imt.record_bound(p)


# ---- errors --------------------------------------------------------------------

def _read_raw(text, pic=None, me=None):
    p = pic or abi.dme_imt_picture()
    n = C.c_size_t()
    return abi.lib().dme_imt_read_picture(text, len(text), C.byref(p), None if me is None else C.byref(me.struct()),
                                          C.byref(n))


@pytest.mark.parametrize("bad", [
    b"[frame:1>mo_comp\n\n1 0 8 8 1 1 1 1",            # missing ']'
    b"\n[frame:1]>p_comp\n",                             # unknown sort
    b"\n[frame:x]>intra\n",                              # picture number
    b"\n[frame:1]>mo_comp\n\n3 0 1 2 8 8 1 1 1 1\n",     # three references
    b"\n[frame:1]>mo_comp\n\n1 0 8 8 4000 4000 16000 16000\n\n0 1 \n",  # arrays longer than the text
])
def test_reader_rejects_malformed_headers(bad):
    with pytest.raises(DmeError):
        imt.read_picture(bad)


def test_reader_rejects_malformed_bodies_and_leaves_arrays_untouched():
    me = _known()
    good = imt.write_picture(7, _geom(2, 3, 1, 2), [3, 11], me)
    cases = [
        good[:-20],                                      # truncated
        good.replace(b"-128 0 127", b"-128 0 32768"),    # DC beyond a short
        good.replace(b"1.5 1.23457e+06", b"1.5 1e+39"),  # beyond a float
        good.replace(b"0 1 3 ", b"0 one 3 "),            # not a number
        good.replace(b"7.125 2", b"7.125 2x"),           # trailing garbage
    ]
    for bad in cases:
        assert bad != good
        with pytest.raises(DmeError):
            imt.read_picture(bad)
        target = _me(2, 3, 1, 2)
        for a in [target.mode, target.sb_split]:
            a[...] = 9
        pic = abi.dme_imt_picture()
        pic.sb_ylen, pic.sb_xlen, pic.mv_ylen, pic.mv_xlen = 1, 2, 2, 3
        assert _read_raw(bad, pic, target) == abi.DME_ERR_INVALID
        assert (target.mode == 9).all() and (target.sb_split == 9).all()
    # the caller's shapes must match the record's
    pic = abi.dme_imt_picture()
    pic.sb_ylen, pic.sb_xlen, pic.mv_ylen, pic.mv_xlen = 1, 2, 3, 2
    assert _read_raw(good, pic, _me(3, 2, 1, 2)) == abi.DME_ERR_INVALID
    assert b"differ" in abi.lib().dme_last_error()


def test_writer_rejects_bad_arguments():
    me = _known()
    g = _geom(2, 3, 1, 2)
    with pytest.raises(DmeError):
        imt.write_picture(1, g, [1, 2, 3], me)
    with pytest.raises(DmeError):
        imt.write_picture(1, g, [], me)
    p = imt.picture_info(1, g, [1, 2])
    m = me.struct()
    m.bipred_costs = None
    n = C.c_size_t()
    assert abi.lib().dme_imt_write_picture(C.byref(p), C.byref(m), None, 0, C.byref(n)) == abi.DME_ERR_INVALID
    m = me.struct()
    assert abi.lib().dme_imt_write_picture(C.byref(p), C.byref(m), None, 0, C.byref(n)) == abi.DME_OK
    buf = C.create_string_buffer(n.value)
    assert abi.lib().dme_imt_write_picture(C.byref(p), C.byref(m), buf, n.value - 1, C.byref(n)) == \
        abi.DME_ERR_INVALID
    assert buf.raw == b"\0" * len(buf.raw)  # nothing written
