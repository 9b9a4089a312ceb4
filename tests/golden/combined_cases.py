"""Combined-component ME cases (EncoderParams::CombinedME): the pictures of
each case, regenerated from its seed, with chroma for the references too
(the combined picture reads every picture's Y, U and V), and the case table
that make_golden.py writes combined.npz from and the tests read.

A "bright" case lifts luma by 64 and scales chroma by 1.5 (both clipped to
8-bit range), so that a large share of combined samples lie above 127 and
need the 9-bit paths."""
import hashlib

import numpy as np

from dirac_me import DME_FP_C, DME_FP_MMX, synth

# (name, seed, size, nrefs, precision, depth (-1: the reference's formula), lambda, fp, variant, bright)
CASES = [
    ("cc_cif2_q", 41, "cif", 2, 2, 3, 10.0, DME_FP_MMX, "mmx", False),
    ("cc_cif1_h", 42, "cif", 1, 1, 3, 10.0, DME_FP_MMX, "mmx", False),
    ("cc_cif2_p", 43, "cif", 2, 0, 3, 10.0, DME_FP_MMX, "mmx", False),
    ("cc_cif2_e", 44, "cif", 2, 3, 3, 10.0, DME_FP_MMX, "mmx", False),
    ("cc_cif2_q_c", 45, "cif", 2, 2, 3, 10.7, DME_FP_C, "c", False),
    ("cc_cif2_dref", 46, "cif", 2, 2, -1, 10.0, DME_FP_MMX, "mmx", False),
    ("cc_cif2_q_bright", 47, "cif", 2, 2, 3, 10.0, DME_FP_MMX, "mmx", True),
    ("cc_cif2_e_bright", 48, "cif", 2, 3, 3, 10.0, DME_FP_MMX, "mmx", True),
    ("cc_720p2_q", 49, "720p", 2, 2, 4, 10.0, DME_FP_MMX, "mmx", True),
]


def _lift(y, u, v):
    y = np.clip(y.astype(np.int32) + 64, -128, 127).astype(np.int16)
    u = np.clip(np.round(u.astype(np.float64) * 1.5), -128, 127).astype(np.int16)
    v = np.clip(np.round(v.astype(np.float64) * 1.5), -128, 127).astype(np.int16)
    return y, u, v


def pictures(seed, size, nrefs, bright=False):
    """d with "cur", "u", "v", "refs" (luma) and "refs_yuv" ((y, u, v) per
    reference), as synth.sized makes them plus reference chroma."""
    d = synth.sized(seed, size, nrefs)
    refs = [(r,) + synth.chroma(r, 1000 + 10 * seed + k) for k, r in enumerate(d["refs"])]
    cur = (d["cur"], d["u"], d["v"])
    if bright:
        cur = _lift(*cur)
        refs = [_lift(*r) for r in refs]
    d["cur"], d["u"], d["v"] = cur
    d["refs_yuv"] = refs
    d["refs"] = [r[0] for r in refs]
    return d


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def inputs_digest(d):
    parts = [d["cur"], d["u"], d["v"]] + [a for r in d["refs_yuv"] for a in r]
    return sha(np.concatenate([p.ravel() for p in parts]))
