"""Pins of the oracle's motion compensation (oracle_motion_compensate, a
restatement of MotionCompensator::CompensatePicture, mot_comp.cpp:63-1230).

1. Against the fixtures written from the reference's own MotionCompensator
   (tests/golden/make_golden.py --mc-only): bit-exact, all three components,
   both directions, every vector precision, padded arrays, one and two
   references, scrambled vectors/modes/splits, and DC values that drive the
   16-bit OBMC sums through wrap-around (MMX and plain-C builds differ there).
2. Against the reference build itself, live, where oracle/_ref exists.
3. The reference's own unit test of this path (unit_tests/motion_comp_test.cpp:
   testRef1and2): adding the ref-1 and the ref-2 predictions one after the
   other comes within 5 of adding the bi-prediction twice, every precision.
   (Its other cases compensate an intra picture, which CompensatePicture
   leaves unchanged, mot_comp.cpp:142: callers compensate inter pictures only.)
4. Against what the mathematics fixes, independently of the oracle's code:
   the OBMC weights of every block arrangement sum to 64 at every sample, so
   a flat reference predicts itself; zero-fraction vectors copy reference
   samples (clamped at the edges); subtracting then adding the same
   prediction gives the picture back; a bi-prediction from two identical
   references equals the single-reference prediction.
"""
import os
import sys

import numpy as np
import pytest

import oracle_py as O
from dirac_me import MEData, abi, geometry, make_params

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
import mc_cases as M  # noqa: E402

SUB, ADD = abi.DME_MC_SUBTRACT, abi.DME_MC_ADD


@pytest.mark.parametrize("name", [c[0] for c in M.CASES])
def test_oracle_mc_matches_reference_fixture(name):
    p, refs, me, planes, meta, z = M.load_case(name)
    O.motion_compensate("oracle", p, refs, me, meta["nrefs"], meta["direction"], planes)
    for c in range(3):
        if name + "/out%d" % c in z.files:
            assert np.array_equal(planes[c], z[name + "/out%d" % c]), c
        assert M.sha(planes[c]) == meta["out_sha256"][c], c


# ---- live comparison with the reference build -----------------------------

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) not present")


@needs_ref
@pytest.mark.parametrize("seed,nrefs,prec,pad,blocks", [
    (60, 2, 2, (0, 0), (12, 12, 8, 8)),
    (61, 1, 3, (24, 4), (16, 12, 12, 8)),
    (62, 2, 1, (8, 8), (8, 8, 8, 8)),
    (63, 2, 0, (0, 12), (24, 24, 16, 16)),
])
@pytest.mark.parametrize("direction", [SUB, ADD])
def test_oracle_mc_matches_reference_live_scrambled(seed, nrefs, prec, pad, blocks, direction):
    d, refs, src = M.pictures(seed, nrefs)
    p = make_params(*blocks, precision=2, depth=d["depth"], lam=10.0)
    me = MEData(geometry(p, 352, 288))
    M.scramble(me, seed, nrefs)
    p.mv_precision = prec
    for variant in ("mmx", "c"):
        want = M.padded(src, pad)
        got = M.padded(src, pad)
        O.motion_compensate(variant, p, refs, me, nrefs, direction, want)
        O.motion_compensate("oracle", p, refs, me, nrefs, direction, got)
        for c in range(3):
            assert np.array_equal(want[c], got[c]), (variant, c)


# ---- closed forms and invariants -------------------------------------------

def _flat_case(nrefs=1, blocks=(12, 12, 8, 8), seed=70):
    d, refs, src = M.pictures(seed, nrefs)
    p = make_params(*blocks, precision=2, depth=d["depth"], lam=10.0)
    me = MEData(geometry(p, 352, 288))
    M.scramble(me, seed, nrefs)  # random splits: every unit shape and edge class
    return p, refs, src, me


@pytest.mark.parametrize("blocks", [(12, 12, 8, 8), (16, 12, 12, 8), (8, 8, 8, 8), (24, 24, 16, 16)])
def test_flat_reference_predicts_itself(blocks):
    p, refs, src, me = _flat_case(blocks=blocks)
    flat = tuple(np.full_like(a, -37) for a in refs[0])
    me.mode[...] = abi.DME_REF1_ONLY  # scrambled vectors (reads clamped at the edges) and splits
    planes = [a.copy() for a in src]
    O.motion_compensate("oracle", p, [flat], me, 1, SUB, planes)
    for c in range(3):
        assert np.array_equal(planes[c], src[c] + 37), c


def test_dc_blocks_predict_their_value():
    p, refs, src, me = _flat_case()
    me.mode[...] = abi.DME_INTRA
    for a in me.dc:
        a[...] = 21
    planes = [a.copy() for a in src]
    O.motion_compensate("oracle", p, refs, me, 1, SUB, planes)
    for c in range(3):
        assert np.array_equal(planes[c], src[c] - 21), c


@pytest.mark.parametrize("dx,dy", [(0, 0), (3, -2), (-7, 5), (6, -4)])
def test_integer_vectors_copy_reference_samples(dx, dy):
    p, refs, src, me = _flat_case()
    me.mode[...] = abi.DME_REF1_ONLY
    me.vectors[0]["x"] = 4 * dx  # quarter-pel units
    me.vectors[0]["y"] = 4 * dy
    planes = [np.zeros_like(a) for a in src]
    O.motion_compensate("oracle", p, refs, me, 1, ADD, planes)
    for c in range(3):
        if c and (dx % 2 or dy % 2):
            continue  # chroma vectors are halved: odd luma displacements land between chroma samples
        ex, ey = (dx // 2, dy // 2) if c else (dx, dy)
        r = refs[0][c]
        h, w = r.shape
        ys = np.clip(np.arange(h) + ey, 0, h - 1)
        xs = np.clip(np.arange(w) + ex, 0, w - 1)
        assert np.array_equal(planes[c], r[np.ix_(ys, xs)]), c


@pytest.mark.parametrize("prec", [0, 1, 2, 3])
def test_subtract_then_add_is_identity(prec):
    p, refs, src, me = _flat_case(nrefs=2, seed=71)
    p.mv_precision = prec
    planes = [a.copy() for a in src]
    O.motion_compensate("oracle", p, refs, me, 2, SUB, planes)
    assert any(not np.array_equal(planes[c], src[c]) for c in range(3))
    O.motion_compensate("oracle", p, refs, me, 2, ADD, planes)
    for c in range(3):
        assert np.array_equal(planes[c], src[c]), c


def test_bipred_of_identical_references_equals_single():
    p, refs, src, me = _flat_case(nrefs=2, seed=72)
    me.vectors[1][...] = me.vectors[0]
    me.mode[...] = np.where(me.mode == abi.DME_INTRA, abi.DME_INTRA, abi.DME_REF1AND2)
    both = [a.copy() for a in src]
    O.motion_compensate("oracle", p, [refs[0], refs[0]], me, 2, SUB, both)
    me.mode[...] = np.where(me.mode == abi.DME_INTRA, abi.DME_INTRA, abi.DME_REF1_ONLY)
    one = [a.copy() for a in src]
    O.motion_compensate("oracle", p, [refs[0]], me, 1, SUB, one)
    for c in range(3):
        assert np.array_equal(both[c], one[c]), c


def test_padding_rules():
    p, refs, src, me = _flat_case()
    planes = M.padded(src, (16, 8))
    O.motion_compensate("oracle", p, refs, me, 1, SUB, planes)
    for c in range(3):
        h, w = src[c].shape
        assert not planes[c][:, w:].any() and not planes[c][h:, :].any(), c
    planes = M.padded(src, (16, 8))
    O.motion_compensate("oracle", p, refs, me, 1, ADD, planes)
    for c in range(3):
        h, w = src[c].shape
        a = planes[c]
        assert np.array_equal(a[:h, w:], np.repeat(a[:h, w - 1:w], a.shape[1] - w, axis=1)), c
        assert np.array_equal(a[h:, :], np.repeat(a[h - 1:h, :], a.shape[0] - h, axis=0)), c


# ---- the reference's unit test (motion_comp_test.cpp:378-460) ---------------

@pytest.mark.parametrize("prec", [0, 1, 2, 3])
@pytest.mark.parametrize("split", [0, 1, 2])
def test_reference_unit_test_ref1_plus_ref2_is_twice_bipred(prec, split):
    refs = [M.ramp_picture(0), M.ramp_picture(50)]
    p = make_params(12, 12, 8, 8, precision=min(prec, 2), depth=2, lam=10.0)
    p.mv_precision = prec
    g = geometry(make_params(12, 12, 8, 8, precision=2, depth=2, lam=10.0), 352, 288)
    zero = [np.zeros((288, 352), np.int16), np.zeros((144, 176), np.int16), np.zeros((144, 176), np.int16)]
    two = [a.copy() for a in zero]
    O.motion_compensate("oracle", p, refs, M.uniform_mv(MEData(g), abi.DME_REF1_ONLY, (5, 5), (5, 5), split), 2, ADD,
                        two)
    O.motion_compensate("oracle", p, refs, M.uniform_mv(MEData(g), abi.DME_REF2_ONLY, (7, 3), (7, 3), split), 2, ADD,
                        two)
    bi = [a.copy() for a in zero]
    me = M.uniform_mv(MEData(g), abi.DME_REF1AND2, (5, 5), (7, 3), split)
    O.motion_compensate("oracle", p, refs, me, 2, ADD, bi)
    O.motion_compensate("oracle", p, refs, me, 2, ADD, bi)
    for c in range(3):
        assert np.abs(two[c].astype(int) - bi[c]).max() <= 5, c
        assert np.abs(two[c]).max() > 0, c
