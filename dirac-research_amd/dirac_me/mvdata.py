"""The MV-data wire format of an inter picture through the C ABI
(dme_mvdata_* in include/dme.h): the bytes PictureCompressor::CodeMVData
writes (libdirac_encoder/picture_compress.cpp:463-530) and
PictureDecompressor::DecompressMVData reads (libdirac_decoder/
picture_decompress.cpp:266-357).
"""
import ctypes as C

from . import MEData, abi
from .abi import check, lib


def info(geom, num_refs, mv_precision, weights=(1, 1, 1)):
    """The header fields of a picture estimated with geometry geom
    (dirac_me.geometry): its level-2 luma block parameters, grid and
    references; the picture weights default to the encoder's."""
    i = abi.dme_mvdata_info()
    i.xnum_sb, i.ynum_sb, i.num_refs = geom.xnum_sb, geom.ynum_sb, num_refs
    i.xblen, i.yblen = geom.luma_blen[2][0], geom.luma_blen[2][1]
    i.xbsep, i.ybsep = geom.luma_bsep[2][0], geom.luma_bsep[2][1]
    i.mv_precision = mv_precision
    i.weights_bits, i.ref1_weight, i.ref2_weight = weights
    return i


def encode(inf, me):
    """The MV data of MEData me (sb_split, mode, vectors, dc) as bytes."""
    m = me.struct()
    n = C.c_size_t()
    check(lib().dme_mvdata_encode(C.byref(inf), C.byref(m), None, 0, C.byref(n)))
    buf = C.create_string_buffer(max(n.value, 1))
    check(lib().dme_mvdata_encode(C.byref(inf), C.byref(m), buf, n.value, C.byref(n)))
    return buf.raw[:n.value]


def encode_batch(infos, mes, threads=0, cap=None):
    """dme_mvdata_encode_batch over pictures; cap bytes per picture (default:
    sized by a single encode of each)."""
    n = len(mes)
    arr = (abi.dme_mvdata_info * n)(*infos)
    structs = [m.struct() for m in mes]
    md = (abi.dme_medata * n)(*structs)
    caps = [cap] * n if cap else [len(encode(i, m)) for i, m in zip(infos, mes)]
    bufs = [C.create_string_buffer(max(c, 1)) for c in caps]
    ptrs = (C.c_void_p * n)(*[C.addressof(b) for b in bufs])
    capa = (C.c_size_t * n)(*caps)
    lens = (C.c_size_t * n)()
    check(lib().dme_mvdata_encode_batch(n, arr, md, ptrs, capa, lens, threads))
    return [b.raw[:l] for b, l in zip(bufs, lens)]


def decode(data, geom, num_refs):
    """(info, MEData, bytes consumed) of MV data for a picture of geometry geom."""
    inf = abi.dme_mvdata_info()
    inf.xnum_sb, inf.ynum_sb, inf.num_refs = geom.xnum_sb, geom.ynum_sb, num_refs
    me = MEData(geom)
    m = me.struct()
    used = C.c_size_t()
    check(lib().dme_mvdata_decode(data, len(data), C.byref(inf), C.byref(m), C.byref(used)))
    return inf, me, used.value
