"""Motion-compensation cases shared by the fixture writer (make_golden.py) and
the tests: the pictures come from the deterministic synthetic generator, the
MvData arrays (vectors, modes, DC values, superblock splits) are stored in the
fixture itself, so a case does not depend on the motion estimator.

A case is (name, seed, nrefs, precision, pad, direction, fp, blocks, scramble,
big_dc, full):
  pad       extra (columns, rows) of luma beyond the picture (chroma half),
            filled with 7 before compensation;
  scramble  replace the estimated MvData by random vectors (up to +-300
            units), random modes and random splits;
  big_dc    all blocks DC-predicted with DC values in [-3000, 3000), which
            drives the 16-bit OBMC sums through their wrap-around;
  full      keep the whole compensated planes in the fixture (else only
            their SHA-256).
"""
import hashlib
import json
import os

import numpy as np

from dirac_me import DME_FP_C, DME_FP_MMX, MEData, abi, geometry, make_params, synth

HERE = os.path.dirname(os.path.abspath(__file__))

SUB, ADD = abi.DME_MC_SUBTRACT, abi.DME_MC_ADD

CASES = [
    ("mc_cif_q_sub", 41, 2, 2, (0, 0), SUB, DME_FP_MMX, None, False, False, True),
    ("mc_cif_q_add", 41, 2, 2, (0, 0), ADD, DME_FP_MMX, None, False, False, False),
    ("mc_cif_h_sub", 41, 2, 1, (0, 0), SUB, DME_FP_MMX, None, False, False, False),
    ("mc_cif_p_add", 41, 2, 0, (0, 0), ADD, DME_FP_MMX, None, False, False, False),
    ("mc_cif_q_1ref_pad_add", 42, 1, 2, (16, 8), ADD, DME_FP_MMX, None, False, False, True),
    ("mc_cif_q_1ref_pad_sub", 42, 1, 2, (16, 8), SUB, DME_FP_MMX, None, False, False, False),
    ("mc_cif_q_blk16_pad_sub", 43, 2, 2, (32, 16), SUB, DME_FP_MMX, (16, 16, 8, 12), False, False, False),
    ("mc_cif_e_scr_pad_add", 47, 2, 3, (8, 8), ADD, DME_FP_MMX, None, True, False, False),
    ("mc_cif_q_scr_sub", 45, 2, 2, (0, 0), SUB, DME_FP_MMX, None, True, False, False),
    ("mc_cif_h_scr_1ref_add", 46, 1, 1, (0, 0), ADD, DME_FP_MMX, None, True, False, False),
    ("mc_cif_dc_wrap_add_mmx", 48, 2, 2, (0, 0), ADD, DME_FP_MMX, None, False, True, False),
    ("mc_cif_dc_wrap_add_c", 48, 2, 2, (0, 0), ADD, DME_FP_C, None, False, True, False),
]

BY_NAME = {c[0]: c for c in CASES}


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def pictures(seed, nrefs):
    """(d, refs, src): the synthetic CIF set, the references as (y, u, v)
    triples and the current picture's (y, u, v)."""
    d = synth.sized(seed, "cif", nrefs)
    refs = []
    for i, r in enumerate(d["refs"]):
        ru, rv = synth.chroma(r, 100 + i)
        refs.append((r, ru, rv))
    return d, refs, [d["cur"], d["u"], d["v"]]


def inputs_digest(refs, src):
    return sha(np.concatenate([a.ravel() for r in refs for a in r] + [a.ravel() for a in src]))


def padded(src, pad):
    """The three arrays to compensate: the picture's planes inside a larger
    array filled with 7 (luma pad, chroma half of it)."""
    out = []
    for c, a in enumerate(src):
        h, w = a.shape
        px, py = (pad[0], pad[1]) if c == 0 else (pad[0] // 2, pad[1] // 2)
        b = np.full((h + py, w + px), 7, np.int16)
        b[:h, :w] = a
        out.append(b)
    return out


def mv_arrays(me, nrefs):
    return dict(vectors1=me.vectors[0], vectors2=me.vectors[1 if nrefs > 1 else 0], mode=me.mode, dc_y=me.dc[0],
                dc_u=me.dc[1], dc_v=me.dc[2], sb_split=me.sb_split)


def load_mv(me, z):
    """Copies the fixture's MvData arrays into a dirac_me.MEData."""
    me.vectors[0][...] = z["vectors1"]
    me.vectors[1][...] = z["vectors2"]
    me.mode[...] = z["mode"]
    for c, k in enumerate(("dc_y", "dc_u", "dc_v")):
        me.dc[c][...] = z[k]
    me.sb_split[...] = z["sb_split"]
    return me


def scramble(me, seed, nrefs):
    rng = np.random.default_rng(seed)
    for v in me.vectors:
        v["x"] = rng.integers(-300, 300, v.shape)
        v["y"] = rng.integers(-300, 300, v.shape)
    me.mode[...] = rng.integers(0, 4 if nrefs == 2 else 2, me.mode.shape)
    for a in me.dc:
        a[...] = rng.integers(-128, 128, a.shape)
    me.sb_split[...] = rng.integers(0, 3, me.sb_split.shape)


def big_dc(me, seed):
    rng = np.random.default_rng(seed)
    me.mode[...] = abi.DME_INTRA
    for a in me.dc:
        a[...] = rng.integers(-3000, 3000, a.shape)


def load_case(name):
    """(params, refs, me, planes, meta, z) of fixture case `name` from
    mc.npz, its pictures regenerated and checked against the fixture's
    digest; planes are the (padded) arrays before compensation."""
    z = np.load(os.path.join(HERE, "mc.npz"))
    meta = json.loads(str(z["meta"]))[name]
    d, refs, src = pictures(meta["seed"], meta["nrefs"])
    assert inputs_digest(refs, src) == meta["inputs_sha256"], "synthetic generator changed: regenerate mc.npz"
    p = make_params(*meta["block"], precision=min(meta["precision"], 2), depth=d["depth"], lam=10.0, fp=meta["fp"])
    me = MEData(geometry(p, 352, 288))
    load_mv(me, {k.split("/")[1]: z[k] for k in z.files if k.startswith(name + "/")})
    p.mv_precision = meta["precision"]
    return p, refs, me, padded(src, tuple(meta["pad"])), meta, z


def ramp_picture(start, width=352, height=288):
    """PicturesTest::setupPicture (unit_tests/frames_test.cpp:49-80): every
    component filled in raster order with start+1, start+2, ... kept to a
    signed char, as the reference's motion-compensation unit tests use."""
    out = []
    for w, h in ((width, height), (width // 2, height // 2), (width // 2, height // 2)):
        k = np.arange(1, w * h + 1, dtype=np.int64) + start
        out.append((((k + 128) % 256) - 128).astype(np.int16).reshape(h, w))
    return tuple(out)


def uniform_mv(me, mode, v1, v2, split):
    """setupMV1Data / setupMV2Data (unit_tests/motion_comp_test.cpp:70-99):
    one mode and one vector per reference for every block; DC 0."""
    me.mode[...] = mode
    me.vectors[0]["x"], me.vectors[0]["y"] = v1
    me.vectors[1]["x"], me.vectors[1]["y"] = v2
    for a in me.dc:
        a[...] = 0
    me.sb_split[...] = split
    return me
