"""The .imt fixture's cases, shared by the fixture writer (make_golden.py
--imt-only) and the tests.

me_cif.imt is one diagnostics file in the reference's format: the sequence
header, an intra record, then one record per CIF case below.  Each case's
pictures come from the deterministic synthetic generator (its own seed), its
MEData from the reference's motion estimator run on them, and the record's
MEData text from the reference's own operator<<(ostream&, MEData&); the
sequence header and each record's first lines follow WriteDiagnosticsHeader /
WriteDiagnosticsData (encoder/encmain.cpp:165-224), which is not buildable
here.  The 1080p case is kept as the SHA-256 and size of its record
(imt.json).
"""
import hashlib
import json
import os

import numpy as np

from dirac_me import make_params, synth

HERE = os.path.dirname(os.path.abspath(__file__))
IMT_FILE = os.path.join(HERE, "me_cif.imt")
INDEX_FILE = os.path.join(HERE, "imt.json")

SEQUENCE = dict(chroma=2, width=352, height=288, source_sampling=0, topfieldfirst=1, frame_rate_num=25,
                frame_rate_den=1, pix_asr_num=12, pix_asr_den=11, picture_coding_mode=0)
INTRA_PIC = 0

# (name, picture number, size, seed, reference picture numbers)
CIF_CASES = [
    ("cif_p1", 1, "cif", 90, [0]),
    ("cif_b2", 2, "cif", 91, [0, 4]),
]
HD_CASE = ("hd_b3", 3, "1080p", 92, [0, 6])


def sha(b):
    return hashlib.sha256(b).hexdigest()


def inputs(case):
    """(d, params, pic_nums) of a case: the synthetic pictures, the ME
    parameters (quarter pel, lambda 10, the size's blocks and depth) and the
    picture numbers."""
    name, num, size, seed, refs = case
    d = synth.sized(seed, size, len(refs))
    p = make_params(*d["block"], precision=2, depth=d["depth"], lam=10.0)
    return d, p, (num, list(refs))


def inputs_digest(d):
    return sha(np.concatenate([d["cur"].ravel()] + [r.ravel() for r in d["refs"]] +
                              [d["u"].ravel(), d["v"].ravel()]).tobytes())


def load():
    """(the fixture's bytes, its index dict)."""
    with open(IMT_FILE, "rb") as f:
        text = f.read()
    with open(INDEX_FILE) as f:
        return text, json.load(f)
