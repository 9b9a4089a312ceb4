"""The .imt diagnostics text of the encoder's MEData through the C ABI
(dme_imt_* in include/dme.h): a sequence header, then one record per coded
picture.

It mirrors the reference's writer, WriteDiagnosticsHeader and
WriteDiagnosticsData (encoder/encmain.cpp:165-330), and its reader,
ReadInstrumentationHeader (util/instrumentation/instrmain.cpp:84-118) and
ProcessSequence::AddPictureEntry (util/instrumentation/process_sequence.cpp:
119-210).  Text is bytes (ASCII).
"""
import ctypes as C
import types

from . import MEData, abi
from .abi import check, lib

SEQUENCE_FIELDS = ("chroma", "width", "height", "source_sampling", "topfieldfirst", "frame_rate_num",
                   "frame_rate_den", "pix_asr_num", "pix_asr_den", "picture_coding_mode")


def _emit(call, bound=None):
    """Runs a writer once into a buffer of `bound` bytes (or sizes it first)."""
    n = C.c_size_t()
    if bound is None:
        check(call(None, 0, C.byref(n)))
        bound = n.value
    buf = C.create_string_buffer(bound)
    check(call(buf, bound, C.byref(n)))
    return buf.raw[:n.value]


def record_bound(p):
    """An upper bound on a record's size: at most 12 characters and a space per
    number ("%.6g" of a float, an int32), a newline per row, 256 bytes for the
    first lines and blank lines."""
    S, B = p.sb_ylen * p.sb_xlen, p.mv_ylen * p.mv_xlen
    values = 2 * S + B * (2 + (2 if p.num_refs > 1 else 0) + 3 + 4 * p.num_refs)
    rows = 2 * p.sb_ylen + p.mv_ylen * (2 + (p.num_refs > 1) + 3 + 2 * p.num_refs)
    return 13 * values + rows + 256


def write_sequence(width, height, chroma=2, source_sampling=0, topfieldfirst=1, frame_rate=(24, 1), pix_asr=(1, 1),
                   picture_coding_mode=0):
    """The sequence header (chroma is a ChromaFormat value: format420 is 2)."""
    s = abi.dme_imt_sequence(chroma, width, height, source_sampling, int(topfieldfirst), frame_rate[0], frame_rate[1],
                             pix_asr[0], pix_asr[1], picture_coding_mode)
    return _emit(lambda b, cap, n: lib().dme_imt_write_sequence(C.byref(s), b, cap, n))


def picture_info(pic_num, geom=None, refs=(), intra=False):
    """The record header of picture pic_num estimated with geometry geom
    (dirac_me.geometry) against refs."""
    p = abi.dme_imt_picture()
    p.pic_num = pic_num
    p.intra = int(intra)
    if not intra:
        p.num_refs = len(refs)
        for i, r in enumerate(refs[:2]):  # the writer rejects num_refs outside 1..2
            p.refs[i] = r
        p.ybsep, p.xbsep = geom.luma_bsep[2][1], geom.luma_bsep[2][0]
        p.sb_ylen, p.sb_xlen = geom.ynum_sb, geom.xnum_sb
        p.mv_ylen, p.mv_xlen = geom.ynum_blocks, geom.xnum_blocks
    return p


def write_picture(pic_num, geom, refs, me):
    """One motion-compensated picture's record from a dirac_me.MEData."""
    p = picture_info(pic_num, geom, refs)
    m = me.struct()
    bound = record_bound(p) if 1 <= p.num_refs <= 2 else None
    return _emit(lambda b, cap, n: lib().dme_imt_write_picture(C.byref(p), C.byref(m), b, cap, n), bound)


def write_intra(pic_num):
    p = picture_info(pic_num, intra=True)
    return _emit(lambda b, cap, n: lib().dme_imt_write_picture(C.byref(p), None, b, cap, n))


def read_sequence(text, offset=0):
    """(dict of the header's fields, offset after it)."""
    s = abi.dme_imt_sequence()
    n = C.c_size_t()
    check(lib().dme_imt_read_sequence(text[offset:], len(text) - offset, C.byref(s), C.byref(n)))
    return {f: getattr(s, f) for f in SEQUENCE_FIELDS}, offset + n.value


def _info(p):
    return dict(pic_num=p.pic_num, intra=bool(p.intra), refs=list(p.refs[:p.num_refs]), ybsep=p.ybsep,
                xbsep=p.xbsep, sb_shape=(p.sb_ylen, p.sb_xlen), mv_shape=(p.mv_ylen, p.mv_xlen))


def read_picture(text, offset=0):
    """(info dict, MEData or None for an intra record, offset after it).
    PredCosts/BiPredCosts totals are not in the text and read as 0."""
    rest = text[offset:]
    p = abi.dme_imt_picture()
    n = C.c_size_t()
    check(lib().dme_imt_read_picture(rest, len(rest), C.byref(p), None, C.byref(n)))
    if p.intra:
        return _info(p), None, offset + n.value
    me = MEData(types.SimpleNamespace(ynum_blocks=p.mv_ylen, xnum_blocks=p.mv_xlen, ynum_sb=p.sb_ylen,
                                      xnum_sb=p.sb_xlen))
    m = me.struct()
    check(lib().dme_imt_read_picture(rest, len(rest), C.byref(p), C.byref(m), C.byref(n)))
    return _info(p), me, offset + n.value


def read(text):
    """(sequence dict, [(info, MEData or None), ...]) of a whole .imt text."""
    seq, off = read_sequence(text)
    out = []
    while off < len(text):
        info, me, off = read_picture(text, off)
        out.append((info, me))
    return seq, out
