"""Writes the golden fixtures under tests/golden/ from the reference encoder's
own ME code, compiled from its sources by oracle/Makefile (target `ref`).

Run in a container that has the reference tree:  python tests/golden/make_golden.py
(--mc-only rewrites only the motion-compensation fixture, mc.npz; --full-only
only the full-search fixture, full_search.npz; --imt-only only the .imt
diagnostics fixture, me_cif.imt and imt.json; --eighth-only only the
eighth-pel fixture, eighth.npz; --combined-only only the combined-component
fixture, combined.npz).

Each fixture holds the inputs' SHA-256 (inputs are regenerated from their
seed by the deterministic generator in libdme.so), the parameters, and the
reference's outputs: full MEData arrays for CIF pictures, per-array SHA-256
digests plus the small arrays for a 1080p picture, pyramid / up-converted
planes of small random pictures, and lambda maps of random vector fields.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "dirac-research_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, HERE)

import oracle_py as O  # noqa: E402
from dirac_me import DME_FP_C, DME_FP_MMX, make_params, synth  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def inputs_digest(d):
    return sha(np.concatenate([d["cur"].ravel()] + [r.ravel() for r in d["refs"]] + [d["u"].ravel(), d["v"].ravel()]))


# (name, config function, seed, precision, depth(-1 = reference formula), lambda, fp, variant, full arrays)
CASES = [
    ("cif_s0_d3_q", "config0", 0, 2, 3, 10.0, DME_FP_MMX, "mmx", True),
    ("cif_s1_d3_q", "config0", 1, 2, 3, 10.0, DME_FP_MMX, "mmx", True),
    ("cif_s2_dref_q", "config0", 2, 2, -1, 10.0, DME_FP_MMX, "mmx", True),
    ("cif_s0_d3_h", "config0", 0, 1, 3, 10.0, DME_FP_MMX, "mmx", True),
    ("cif_s0_d3_p", "config0", 0, 0, 3, 10.0, DME_FP_MMX, "mmx", True),
    ("cif_s3_d3_q_l10.7", "config0", 3, 2, 3, 10.7, DME_FP_MMX, "mmx", True),
    ("cif_s3_d3_q_l10.7_c", "config0", 3, 2, 3, 10.7, DME_FP_C, "c", True),
    ("hd_s0_d4_q", "config1", 0, 2, 4, 10.0, DME_FP_MMX, "mmx", False),
    ("hd_s0_d4_q_l10.7_c", "config1", 0, 2, 4, 10.7, DME_FP_C, "c", False),
]


def main():
    out_dir = HERE
    index = {}
    for name, cfg, seed, prec, depth, lam, fp, variant, full in CASES:
        d = getattr(synth, cfg)(seed)
        p = make_params(*d["block"], precision=prec, depth=depth, lam=lam, fp=fp)
        res, cap, g = O.run_ref(p, d["cur"], d["refs"], d["u"], d["v"], capture=depth > 0, variant=variant)
        arrs = res.arrays(d["nrefs"])
        meta = dict(config=cfg, seed=seed, precision=prec, depth=depth, lam=lam, fp=fp, variant=variant,
                    nrefs=d["nrefs"], inputs_sha256=inputs_digest(d), sha256={k: sha(v) for k, v in arrs.items()})
        save = {}
        if full:
            save.update(arrs)
            if cap is not None:
                for r in range(d["nrefs"]):
                    for l in range(1, g.depth + 1):
                        save["level%d_vectors%d" % (l, r + 1)] = cap.level_vectors[r][l]
        else:
            for k in ("mode", "sb_split", "sb_costs"):
                save[k] = arrs[k]
        np.savez_compressed(os.path.join(out_dir, name + ".npz"), meta=json.dumps(meta), **save)
        index[name] = meta
        print("wrote", name)

    # pyramid and up-converted planes of small random pictures
    rng = np.random.default_rng(20261203)
    planes = {}
    for i, (w, h) in enumerate(((64, 48), (50, 38), (96, 70))):
        a = rng.integers(-128, 128, size=(h, w)).astype(np.int16)
        planes["in%d" % i] = a
        planes["down%d" % i] = O.downconvert(O.ref(), a, w // 2, h // 2, "refh_downconvert")
        planes["up%d" % i] = O.upconvert(O.ref(), a, "refh_upconvert")
    np.savez_compressed(os.path.join(out_dir, "planes.npz"), **planes)
    print("wrote planes")

    # lambda maps of random vector fields, one and two references
    lam = {}
    for i, (X, Y) in enumerate(((44, 36), (160, 92))):
        for r in range(2):
            v = np.zeros((Y, X), O.abi.MV_DTYPE)
            v["x"] = rng.integers(-6, 7, size=(Y, X))
            v["y"] = rng.integers(-6, 7, size=(Y, X))
            # a few smooth regions and outliers
            v["x"][: Y // 2, : X // 2] = 3
            v["y"][Y // 3:, X // 3:] = -2
            lam["v%d_%d" % (i, r)] = v
        for nrefs in (1, 2):
            lam["lam%d_n%d" % (i, nrefs)] = O.lambda_map(O.ref(), "refh_lambda_map", lam["v%d_0" % i],
                                                         lam["v%d_1" % i], 10.7, nrefs)
    np.savez_compressed(os.path.join(out_dir, "lambda.npz"), **lam)
    print("wrote lambda")
    with open(os.path.join(out_dir, "index.json"), "w") as f:
        json.dump(index, f, indent=1, sort_keys=True)


# full search: (name, seed, nrefs, x range, y range, picture numbers (cur, refs), precision)
FULL_CASES = [
    ("full_cif_s70_r6x5", 70, 2, 6, 5, (7, [4, 9]), 2),   # tdiff 3 and 2
    ("full_cif_s71_r32", 71, 1, 32, 32, (1, [0]), 2),      # the encoder default range
]


def main_full():
    """Full-search fixtures (full_search.npz): the reference's PixelMatcher
    with FullSearch set, then the rest of the live sequence; full MEData."""
    out, index = {}, {}
    for name, seed, nrefs, xr, yr, nums, prec in FULL_CASES:
        d = synth.sized(seed, "cif", nrefs)
        p = make_params(*d["block"], precision=prec, lam=10.0, full_search=True, x_range=xr, y_range=yr)
        res, _, _ = O.run_ref(p, d["cur"], d["refs"], d["u"], d["v"], pic_nums=nums)
        index[name] = dict(seed=seed, nrefs=nrefs, x_range=xr, y_range=yr, pic_nums=[nums[0], list(nums[1])],
                           precision=prec, inputs_sha256=inputs_digest(d))
        for k, a in res.arrays(nrefs).items():
            out[name + "/" + k] = a
        print("wrote", name)
    np.savez_compressed(os.path.join(HERE, "full_search.npz"), meta=json.dumps(index), **out)


# eighth-pel estimation: (name, seed, size, nrefs, depth, lambda, fp, variant);
# two-reference cases exercise BiBlockEighthPel's bounds-checked read of the
# second reference (me_utils.cpp:1764-1768)
EIGHTH_CASES = [
    ("e_cif1_s0", 0, "cif", 1, 3, 10.0, DME_FP_MMX, "mmx"),
    ("e_cif2_s3", 3, "cif", 2, 3, 10.0, DME_FP_MMX, "mmx"),
    ("e_cif2_s7_l10.7", 7, "cif", 2, 3, 10.7, DME_FP_MMX, "mmx"),
    ("e_cif2_s3_c", 3, "cif", 2, 3, 10.0, DME_FP_C, "c"),
    ("e_720p2_s5", 5, "720p", 2, 4, 10.0, DME_FP_MMX, "mmx"),
]


def main_eighth():
    """Eighth-pel fixtures (eighth.npz): the reference's live sequence at
    MV_PRECISION_EIGHTH_PIXEL (BlockDiffEighthPel, BiBlockEighthPel); full
    MEData of the CIF cases, digests plus mode and superblock arrays of the
    720p case."""
    out, index = {}, {}
    for name, seed, size, nrefs, depth, lam, fp, variant in EIGHTH_CASES:
        d = synth.sized(seed, size, nrefs)
        p = make_params(*d["block"], precision=3, depth=depth, lam=lam, fp=fp)
        res, _, _ = O.run_ref(p, d["cur"], d["refs"], d["u"], d["v"], variant=variant)
        arrs = res.arrays(nrefs)
        index[name] = dict(seed=seed, size=size, nrefs=nrefs, depth=depth, lam=lam, fp=fp, variant=variant,
                           inputs_sha256=inputs_digest(d), sha256={k: sha(v) for k, v in arrs.items()})
        for k, a in arrs.items():
            if size == "cif" or k in ("mode", "sb_split", "sb_costs"):
                out[name + "/" + k] = a
        print("wrote", name)
    np.savez_compressed(os.path.join(HERE, "eighth.npz"), meta=json.dumps(index), **out)


def main_combined():
    """Combined-component fixtures (combined.npz, cases in combined_cases.py):
    the reference's live sequence with EncoderParams::CombinedME set, which
    makes every stage read EncPicture::CombinedData / UpCombinedData; full
    MEData of the CIF cases, digests plus mode and superblock arrays of the
    720p case; and the combined pictures of the CIF cases' current
    pictures (EncPicture::CombinedData itself)."""
    import combined_cases as CC

    out, index = {}, {}
    for name, seed, size, nrefs, prec, depth, lam, fp, variant, bright in CC.CASES:
        d = CC.pictures(seed, size, nrefs, bright)
        p = make_params(*d["block"], precision=prec, depth=depth, lam=lam, fp=fp)
        p.combined_me = 1
        res, _, _ = O.run_ref(p, d["cur"], d["refs_yuv"], d["u"], d["v"], variant=variant)
        arrs = res.arrays(nrefs)
        comb = O.combine(d["cur"], d["u"], d["v"], "ref", variant)
        index[name] = dict(seed=seed, size=size, nrefs=nrefs, precision=prec, depth=depth, lam=lam, fp=fp,
                           variant=variant, bright=bright, inputs_sha256=CC.inputs_digest(d),
                           sha256={k: sha(v) for k, v in arrs.items()}, combined_sha256=sha(comb),
                           above_127=float((comb > 127).mean()))
        for k, a in arrs.items():
            if size == "cif" or k in ("mode", "sb_split", "sb_costs"):
                out[name + "/" + k] = a
        if size == "cif":
            out[name + "/combined"] = comb
        print("wrote", name, "combined samples above 127: %.3f" % index[name]["above_127"])
    np.savez_compressed(os.path.join(HERE, "combined.npz"), meta=json.dumps(index), **out)


def main_mc():
    """Motion-compensation fixtures (mc.npz): for each case in mc_cases.CASES,
    the MvData from the reference's own motion estimator (or scrambled), and
    the planes the reference's MotionCompensator::CompensatePicture leaves
    (the build of the case's float semantics: MMX or plain C)."""
    import mc_cases as M
    from dirac_me import MEData, geometry

    out = {}
    index = {}
    for name, seed, nrefs, prec, pad, direction, fp, blocks, scr, bigdc, full in M.CASES:
        d, refs, src = M.pictures(seed, nrefs)
        blk = blocks or d["block"]
        p = make_params(*blk, precision=min(prec, 2), depth=d["depth"], lam=10.0, fp=fp)
        res, _, _ = O.run_ref(p, d["cur"], d["refs"], d["u"], d["v"], variant="mmx")
        me = MEData(geometry(p, d["cur"].shape[1], d["cur"].shape[0]))
        M.load_mv(me, M.mv_arrays(res, nrefs))
        if scr:
            M.scramble(me, seed, nrefs)
        if bigdc:
            M.big_dc(me, seed)
        p.mv_precision = prec
        planes = M.padded(src, pad)
        O.motion_compensate("mmx" if fp == DME_FP_MMX else "c", p, refs, me, nrefs, direction, planes)
        meta = dict(seed=seed, nrefs=nrefs, precision=prec, pad=list(pad), direction=direction, fp=fp,
                    block=list(blk), inputs_sha256=M.inputs_digest(refs, src),
                    out_sha256=[M.sha(a) for a in planes])
        index[name] = meta
        for k, a in M.mv_arrays(me, nrefs).items():
            out[name + "/" + k] = a
        if full:
            for c, a in enumerate(planes):
                out[name + "/out%d" % c] = a
        print("wrote", name)
    np.savez_compressed(os.path.join(HERE, "mc.npz"), meta=json.dumps(index), **out)


def main_imt():
    """The .imt fixture (me_cif.imt, imt.json): the reference's estimates of
    the imt_cases pictures written as a diagnostics file, the MEData text of
    every record by the reference's own operator<<."""
    import imt_cases as I

    text = O.imt_sequence(I.SEQUENCE) + O.imt_intra(I.INTRA_PIC)
    index = dict(records={})
    for case in I.CIF_CASES + [I.HD_CASE]:
        d, p, nums = I.inputs(case)
        res, _, g = O.run_ref(p, d["cur"], d["refs"], d["u"], d["v"], pic_nums=nums)
        rec = O.imt_record_head(nums[0], nums[1], g.luma_bsep[2][1], g.luma_bsep[2][0], res.sb_split.shape,
                                res.mode.shape) + O.ref_imt_medata_text(res, len(nums[1]))
        index["records"][case[0]] = dict(inputs_sha256=I.inputs_digest(d), record_sha256=I.sha(rec),
                                         record_bytes=len(rec))
        if case in I.CIF_CASES:
            text += rec
        print("wrote", case[0], len(rec), "bytes")
    index["file_sha256"] = I.sha(text)
    with open(I.IMT_FILE, "wb") as f:
        f.write(text)
    with open(I.INDEX_FILE, "w") as f:
        json.dump(index, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    if "--full-only" in sys.argv:
        main_full()
    elif "--imt-only" in sys.argv:
        main_imt()
    elif "--mc-only" in sys.argv:
        main_mc()
    elif "--eighth-only" in sys.argv:
        main_eighth()
    elif "--combined-only" in sys.argv:
        main_combined()
    else:
        main()
        main_mc()
        main_full()
        main_imt()
        main_eighth()
        main_combined()
