"""The MV-data wire format (dme_mvdata_*): the bytes PictureCompressor::
CodeMVData writes (libdirac_encoder/picture_compress.cpp:463-530).

The reference's coder cannot be compiled here (its arith_codec.h needs the
configure-generated dirac-stdint.h) and the reference holds no coded
bitstreams, so the byte stream is "parity unpinned".  What pins it:
  * known answers derived by hand from the coder's definition (exp-Golomb
    codes of ByteIO, one-symbol arithmetic-coded elements, the header);
  * the reference's own prediction helpers (GetUMean, GetSMean, Median,
    BlockParametersIndex, SetDefaultBlockParameters) compiled from its
    sources, against the oracle's restatement;
  * the library against the independent Python restatement
    (oracle/mvcodec_ref.py) on MEData written by the reference itself
    (tests/golden) and on random fields;
  * decode(encode(x)) == x, and malformed input refused.
"""
import ctypes as C

import numpy as np
import pytest

import mvcodec_ref as R
import oracle_py as O
import dirac_me
from dirac_me import abi, mvdata


def test_exp_golomb_known_answers():
    # interleaved exp-Golomb of ByteIO::WriteUint (byteio.cpp): value + 1 =
    # 1 b1 b0 ... is written 0 b1 0 b0 ... 1
    want = {0: "1", 1: "001", 2: "011", 3: "00001", 4: "00011", 5: "01001", 6: "01011", 7: "0000001"}
    for v, bits in want.items():
        o = R.ByteOut()
        o.uint(v)
        got = "".join(str((byte >> (7 - i)) & 1) for byte in (o.data + bytes([o.byte])) for i in range(8))
        assert got[:len(bits)] == bits, v
        assert R.ByteIn(bytes(o.data) + bytes([o.byte])).uint() == v
    # WriteSint: magnitude, then a sign bit 1 for negative, 0 for positive
    o = R.ByteOut()
    o.sint(-2)
    o.align()
    assert o.data[0] >> 4 == 0b0111


def test_arithmetic_coder_known_answers():
    # One symbol 0 in a fresh context: p0 = 0x8000, the split of range 0xffff
    # is 0x7fff, range becomes 0x7fff (no renormalisation); the flush emits
    # the top bit of low (0) while low + range - 1 keeps it, then bit 14 of
    # low (0) and its complement (1): 001, padded: 0x20.
    e = R.ArithEncoder()
    e.symbol(0, 0)
    assert e.flush() == b"\x20"
    # One symbol 1: low = 0x7fff, range = 0x8000; the flush writes bit 14 of
    # low (1) and its complement (0): 10, padded: 0x80.
    e = R.ArithEncoder()
    e.symbol(1, 0)
    assert e.flush() == b"\x80"
    # the adaptation: after a 0, p0 grows by LUT[255 - 128] = 1920
    e = R.ArithEncoder()
    e.symbol(0, 3)
    assert e.prob[3] == 0x8000 + 1920 and e.prob[0] == 0x8000
    e.symbol(1, 3)
    assert e.prob[3] == 0x8000 + 1920 - R.LUT[(0x8000 + 1920) >> 8]


def test_arithmetic_coder_round_trips_and_compresses():
    rng = np.random.default_rng(3)
    for n, p1 in [(1, 0.5), (50, 0.1), (2000, 0.02), (2000, 0.5), (3000, 0.97)]:
        syms = (rng.random(n) < p1).astype(int)
        ctxs = rng.integers(0, R.NCTX, n)
        e = R.ArithEncoder()
        for s, c in zip(syms, ctxs):
            e.symbol(s, c)
        data = e.flush()
        d = R.ArithDecoder(data)
        assert [int(d.symbol(c)) for c in ctxs] == list(syms)
        if n >= 2000 and p1 in (0.02, 0.97):
            # skewed symbols cost well under a bit each (the contexts adapt;
            # with 16 contexts learning apart, about twice the entropy here)
            h = -(p1 * np.log2(p1) + (1 - p1) * np.log2(1 - p1))
            assert len(data) * 8 < min(0.35 * n, 3 * h * n)
    # unsigned and signed binarisations
    vals = [0, 1, 2, 3, 7, 8, 100, 1000, 65535, 1 << 20]
    e = R.ArithEncoder()
    for v in vals:
        e.uint(v, 4, 8)
        e.sint(-v, 4, 8)
        e.sint(v, 0, 1)
    d = R.ArithDecoder(e.flush())
    for v in vals:
        assert d.uint(4, 8) == v and d.sint(4, 8) == -v and d.sint(0, 1) == v


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) missing")
def test_prediction_helpers_match_the_reference():
    ref = O.ref()
    ref.refh_umean.restype = C.c_uint
    ref.refh_umean.argtypes = [C.c_void_p, C.c_int]
    ref.refh_smean.argtypes = [C.c_void_p, C.c_int]
    ref.refh_median_list.argtypes = [C.c_void_p, C.c_int]
    rng = np.random.default_rng(5)
    for n in (1, 2, 3):
        for _ in range(400):
            u = rng.integers(0, 3, n).astype(np.uint32)
            assert ref.refh_umean(u.ctypes.data, n) == R.umean(u.tolist())
            s = rng.integers(-300, 300, n).astype(np.int32)
            assert ref.refh_smean(s.ctypes.data, n) == R.smean(s.tolist()), s
            assert ref.refh_median_list(s.ctypes.data, n) == R.median(s.tolist()), s
    # every negative-sum case of three small values (GetSMean's odd branch)
    for a in range(-4, 2):
        for b in range(-4, 2):
            for c in range(-4, 2):
                s = np.array([a, b, c], np.int32)
                assert ref.refh_smean(s.ctypes.data, 3) == R.smean([a, b, c])
                assert ref.refh_smean(s.ctypes.data, 2) == R.smean([a, b])
    for bp in [(8, 8, 4, 4), (12, 12, 8, 8), (16, 16, 12, 12), (24, 24, 16, 16), (16, 16, 8, 8), (12, 16, 8, 12),
               (10, 12, 8, 8)]:
        assert ref.refh_block_params_index(*bp) == R.block_params_index(bp)
    out = (C.c_int * 4)()
    for i in range(1, 5):
        assert ref.refh_default_block_params(i, out) == 0
        assert tuple(out) == R.PRESETS[i - 1]


def _lists(me, nrefs):
    split = me.sb_split.tolist()
    mode = me.mode.tolist()
    vec = [[[(int(v["x"]), int(v["y"])) for v in row] for row in me.vectors[r]] for r in range(2)]
    dc = [me.dc[c].tolist() for c in range(3)]
    return split, mode, vec, dc


def _canonical(me, nrefs):
    """The arrays as the decoder rebuilds them: each unit's top-left values in
    all its blocks, vectors only where the mode uses the reference, DC only
    on intra blocks."""
    out = dirac_me.MEData.__new__(dirac_me.MEData)
    out.sb_split = me.sb_split.copy()
    out.mode = me.mode.copy()
    out.vectors = [np.zeros_like(me.vectors[0]) for _ in range(2)]
    out.dc = [np.zeros_like(me.dc[0]) for _ in range(3)]
    for x, y, step in R.units(me.sb_split.tolist()):
        m = me.mode[y, x]
        out.mode[y:y + step, x:x + step] = m
        for r in range(nrefs):
            if m & (r + 1):
                out.vectors[r][y:y + step, x:x + step] = me.vectors[r][y, x]
        if m == 0:
            for c in range(3):
                out.dc[c][y:y + step, x:x + step] = me.dc[c][y, x]
    return out


def _same(a, b, nrefs):
    assert np.array_equal(a.sb_split, b.sb_split)
    assert np.array_equal(a.mode, b.mode)
    for r in range(nrefs):
        assert a.vectors[r].tobytes() == b.vectors[r].tobytes(), "vectors %d" % (r + 1)
    for c in range(3):
        assert np.array_equal(a.dc[c], b.dc[c]), "dc %d" % c


class _Geom:
    def __init__(self, xsb, ysb, blen=(12, 12), bsep=(8, 8)):
        self.xnum_sb, self.ynum_sb = xsb, ysb
        self.xnum_blocks, self.ynum_blocks = 4 * xsb, 4 * ysb
        self.luma_blen = [None, None, blen]
        self.luma_bsep = [None, None, bsep]


def _random_me(rng, xsb, ysb, nrefs, vrange=40, dcrange=128):
    g = _Geom(xsb, ysb)
    me = dirac_me.MEData.__new__(dirac_me.MEData)
    me.sb_split = rng.integers(0, 3, (ysb, xsb)).astype(np.int32)
    Y, X = 4 * ysb, 4 * xsb
    me.mode = np.zeros((Y, X), np.int32)
    me.vectors = [np.zeros((Y, X), abi.MV_DTYPE) for _ in range(2)]
    me.dc = [np.zeros((Y, X), np.int16) for _ in range(3)]
    nmodes = 4 if nrefs == 2 else 2
    for x, y, step in R.units(me.sb_split.tolist()):
        m = int(rng.integers(0, nmodes))
        if step > 1 and m == 0 and rng.random() < 0.7:
            m = 1  # intra is rare in merged units
        me.mode[y:y + step, x:x + step] = m
        for r in range(2):
            me.vectors[r]["x"][y:y + step, x:x + step] = rng.integers(-vrange, vrange + 1)
            me.vectors[r]["y"][y:y + step, x:x + step] = rng.integers(-vrange, vrange + 1)
        for c in range(3):
            me.dc[c][y:y + step, x:x + step] = rng.integers(-dcrange, dcrange)
    me.intra_costs = np.zeros((Y, X), np.float32)
    me.bipred_costs = np.zeros((Y, X), abi.COST_DTYPE)
    me.sb_costs = np.zeros((ysb, xsb), np.float32)
    me.lambda_map = np.zeros((Y, X), np.float32)
    me.intra_block_ratio = np.zeros(1, np.float32)
    me.pred_costs = [np.zeros((Y, X), abi.COST_DTYPE) for _ in range(2)]  # unused by the codec
    return g, me


def _check(g, me, nrefs, prec, weights=(1, 1, 1)):
    inf = mvdata.info(g, nrefs, prec, weights)
    got = mvdata.encode(inf, me)
    split, mode, vec, dc = _lists(me, nrefs)
    bp = tuple(g.luma_blen[2]) + tuple(g.luma_bsep[2])
    want = R.encode(split, mode, vec, dc, nrefs, bp, prec, weights)
    assert got == want
    inf2, back, used = mvdata.decode(got, g, nrefs)
    assert used == len(got)
    assert (inf2.xblen, inf2.yblen, inf2.xbsep, inf2.ybsep, inf2.mv_precision) == bp + (prec,)
    assert (inf2.weights_bits, inf2.ref1_weight) == weights[:2]
    _same(back, _canonical(me, nrefs), nrefs)
    head, s2, m2, v2, d2 = R.decode(got, g.xnum_sb, g.ynum_sb, nrefs)
    assert head["consumed"] == len(got)
    assert s2 == back.sb_split.tolist() and m2 == back.mode.tolist()
    return got


@pytest.mark.parametrize("name", ["cif_s0_d3_q", "cif_s0_d3_h", "cif_s0_d3_p", "cif_s2_dref_q",
                                  "cif_s3_d3_q_l10.7_c"])
def test_reference_estimates_code_as_the_oracle_does(name):
    # MEData written by the reference's own ME (tests/golden): coded by the
    # library and by the restatement, bytes equal, decoded back
    z = np.load("tests/golden/%s.npz" % name)
    meta = z["meta"]
    me = dirac_me.MEData.__new__(dirac_me.MEData)
    me.sb_split = z["sb_split"].astype(np.int32)
    me.mode = z["mode"].astype(np.int32)
    nrefs = 2 if "vectors2" in z else 1
    me.vectors = [z["vectors1"].view(abi.MV_DTYPE).reshape(me.mode.shape),
                  (z["vectors2"] if nrefs == 2 else z["vectors1"]).view(abi.MV_DTYPE).reshape(me.mode.shape)]
    me.dc = [z["dc_y"].astype(np.int16), z["dc_u"].astype(np.int16), z["dc_v"].astype(np.int16)]
    Y, X = me.mode.shape
    me.pred_costs = [np.zeros((Y, X), abi.COST_DTYPE) for _ in range(2)]
    me.intra_costs = np.zeros((Y, X), np.float32)
    me.bipred_costs = np.zeros((Y, X), abi.COST_DTYPE)
    me.sb_costs = np.zeros(me.sb_split.shape, np.float32)
    me.lambda_map = np.zeros((Y, X), np.float32)
    me.intra_block_ratio = np.zeros(1, np.float32)
    prec = {"_p": 0, "_h": 1}.get(name[-2:], 2)
    g = _Geom(me.sb_split.shape[1], me.sb_split.shape[0])
    data = _check(g, me, nrefs, prec)
    assert 0 < len(data) < 4 * me.mode.size


@pytest.mark.parametrize("seed,nrefs,xsb,ysb", [(0, 1, 3, 2), (1, 2, 5, 4), (2, 2, 11, 9), (3, 1, 1, 1), (4, 2, 1, 7)])
def test_random_fields_code_as_the_oracle_does(seed, nrefs, xsb, ysb):
    rng = np.random.default_rng(seed)
    g, me = _random_me(rng, xsb, ysb, nrefs)
    _check(g, me, nrefs, int(rng.integers(0, 4)))
    # large values: long binarisations, 16-bit DC wrap of the prediction
    g, me = _random_me(rng, xsb, ysb, nrefs, vrange=5000, dcrange=32767)
    _check(g, me, nrefs, 3, weights=(2, -3, 5))


def test_custom_block_parameters_and_weights_in_the_header():
    rng = np.random.default_rng(9)
    g, me = _random_me(rng, 2, 2, 2)
    g.luma_blen, g.luma_bsep = [None, None, (10, 14)], [None, None, (8, 12)]
    data = _check(g, me, 2, 1, weights=(3, 2, -1))
    # index 0 (custom): 1, then 10, 14, 8, 12, precision 1, global motion 0,
    # frame prediction mode 0, weights flag 1 ...
    o = R.ByteOut()
    for v in (0, 10, 14, 8, 12, 1):
        o.uint(v)
    o.bit(0)
    o.uint(0)
    o.bit(1)
    assert data[:len(o.data)] == bytes(o.data)


def test_batch_equals_single_pictures():
    rng = np.random.default_rng(11)
    pics = [_random_me(rng, 6, 5, 1 + (i & 1)) for i in range(9)]
    infos = [mvdata.info(g, 1 + (i & 1), 2) for i, (g, me) in enumerate(pics)]
    single = [mvdata.encode(inf, me) for inf, (g, me) in zip(infos, pics)]
    assert mvdata.encode_batch(infos, [me for g, me in pics], threads=4) == single


def test_malformed_data_is_refused():
    rng = np.random.default_rng(12)
    g, me = _random_me(rng, 4, 3, 2)
    data = mvdata.encode(mvdata.info(g, 2, 2), me)
    ok = 0
    for cut in range(0, len(data) - 1, max(1, len(data) // 40)):
        try:
            mvdata.decode(data[:cut], g, 2)
            ok += 1
        except RuntimeError:
            pass
    assert ok == 0  # every truncation misses an element's bytes
    bad = bytearray(data)
    bad[0] = 0b00000110  # block-parameter index 5..: out of range
    with pytest.raises(RuntimeError, match="index"):
        mvdata.decode(bytes(bad), g, 2)
    # random bytes: refused or decoded into legal fields, never out of range
    for s in range(200):
        junk = np.random.default_rng(s).integers(0, 256, 60).astype(np.uint8).tobytes()
        try:
            _, back, _ = mvdata.decode(junk, g, 2)
        except RuntimeError:
            continue
        assert back.sb_split.min() >= 0 and back.sb_split.max() <= 2
        assert back.mode.min() >= 0 and back.mode.max() <= 3
    # an encoder input outside the modes of one reference
    me.mode[0, 0] = 3
    with pytest.raises(RuntimeError, match="mode"):
        mvdata.encode(mvdata.info(g, 1, 2), me)
