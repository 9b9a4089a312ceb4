"""The C ABI (include/dme.h, include/dme_learned.h) without a GPU: the library loads, exports every
declared symbol, and its host-side parts (geometry, lambda map, synthetic
generator, argument checking) agree with the oracle."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_py as O
import dirac_me
from dirac_me import abi, make_params, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in ("dme.h", "dme_learned.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(dme_[a-z_0-9]+)\s*\(", src))
    return sorted(names)


def test_header_symbols_all_exported():
    names = _declared()
    assert len(names) >= 20
    lib = C.CDLL(abi.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(abi.SIGNATURES), "ctypes mirror out of date"


def test_version():
    assert dirac_me.lib().dme_version() == 100


@pytest.mark.parametrize("W,H", [(352, 288), (1280, 720), (1920, 1080), (3840, 2160), (360, 240)])
@pytest.mark.parametrize("blk", [(12, 12, 8, 8), (16, 16, 12, 12), (24, 24, 16, 16), (10, 14, 7, 9)])
@pytest.mark.parametrize("depth", [-1, 2])
def test_geometry_matches_oracle(W, H, blk, depth):
    p = make_params(*blk, depth=depth)
    a = dirac_me.geometry(p, W, H)
    b = O.geometry(p, W, H)
    assert bytes(a) == bytes(b)


def test_geometry_rejects_bad_params():
    with pytest.raises(dirac_me.DmeError) as e:
        dirac_me.geometry(make_params(precision=4), 352, 288)
    assert e.value.code == abi.DME_ERR_INVALID
    dirac_me.geometry(make_params(precision=3), 352, 288)  # eighth pel
    with pytest.raises(dirac_me.DmeError):
        dirac_me.geometry(make_params(depth=7), 352, 288)  # levels below 12 samples
    with pytest.raises(dirac_me.DmeError):
        dirac_me.geometry(make_params(), 351, 288)


def test_combined_component_me_flag_is_validated():
    # EncoderParams::CombinedME (common.h:1317) is a bool: 0 or 1
    p = make_params()
    p.combined_me = 2
    with pytest.raises(dirac_me.DmeError) as e:
        dirac_me.geometry(p, 352, 288)
    assert e.value.code == abi.DME_ERR_INVALID
    assert "combined" in str(e.value)
    for v in (0, 1):
        p.combined_me = v
        dirac_me.geometry(p, 352, 288)


@pytest.mark.parametrize("nrefs", [1, 2])
@pytest.mark.parametrize("lam", [10.0, 10.7, 3.3])
def test_host_lambda_map_matches_oracle(nrefs, lam):
    rng = np.random.default_rng(nrefs * 100 + int(lam * 10))
    Y, X = 92, 160
    v = [np.zeros((Y, X), abi.MV_DTYPE) for _ in range(2)]
    for r in range(2):
        v[r]["x"] = rng.integers(-4, 5, size=(Y, X))
        v[r]["y"] = rng.integers(-4, 5, size=(Y, X))
        v[r]["x"][10:40, 20:90] = 7
    got = dirac_me.lambda_map_host(v[0], v[1] if nrefs > 1 else None, lam, nrefs)
    want = O.lambda_map(O.oracle(), "oracle_lambda_map", v[0], v[1], lam, nrefs)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_host_lambda_map_matches_reference_fixture():
    z = np.load(os.path.join(ROOT, "tests", "golden", "lambda.npz"))
    for i in range(2):
        for n in (1, 2):
            got = dirac_me.lambda_map_host(z["v%d_0" % i], z["v%d_1" % i], 10.7, n)
            assert np.array_equal(got.view(np.uint32), z["lam%d_n%d" % (i, n)].view(np.uint32))


def test_synth_is_deterministic_and_in_range():
    a = synth.config0(3)
    b = synth.config0(3)
    assert np.array_equal(a["cur"], b["cur"]) and np.array_equal(a["refs"][0], b["refs"][0])
    c = synth.config0(4)
    assert not np.array_equal(a["cur"], c["cur"])
    for x in (a["cur"], a["refs"][0], a["u"], a["v"]):
        assert x.dtype == np.int16 and x.min() >= -128 and x.max() <= 127
    assert a["cur"].std() > 20  # textured


def test_synth_warp_integer_shift():
    cur = synth.texture(1, 64, 48, 0)
    ref = synth.warp(cur, 1.0, (2.0, -1.0))
    # ref(x, y) = cur(x + 2, y - 1) away from the clamped borders
    assert np.array_equal(ref[1:, :-2], cur[:-1, 2:])


def test_create_fails_cleanly_without_device():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    rc = dirac_me.lib().dme_create(0, C.byref(h))
    assert rc == abi.DME_ERR_NO_DEVICE
    assert b"device" in dirac_me.lib().dme_last_error()
