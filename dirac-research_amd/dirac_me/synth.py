"""Seeded synthetic pictures for the five motion-estimation workloads.

Textures: uniform noise smoothed by a Gaussian (sigma 1.5) plus flat
rectangles with noise; references: the picture warped by a motion field of
affine zoom/pan plus moving objects, quantised to quarter pel, plus noise.
All 8-bit values are stored minus 128 (libdirac_common/pic_io.cpp:396-401).
Generation runs in libdme.so's host code (SplitMix64), so pictures are
identical on every machine for a given seed.
"""
import ctypes as C

import numpy as np

from . import abi
from .abi import check, lib, ptr


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def texture(seed, width, height, nrects=3):
    out = np.empty((height, width), np.int16)
    check(lib().dme_synth_texture(seed, width, height, nrects, ptr(out)))
    return out


def warp(cur, zoom=1.0, pan=(0.0, 0.0), objects=(), noise_sigma=0.0, noise_seed=0):
    cur = np.ascontiguousarray(cur, np.int16)
    H, W = cur.shape
    objs = (abi.dme_synth_object * max(1, len(objects)))()
    for i, o in enumerate(objects):
        objs[i] = abi.dme_synth_object(*[float(v) for v in o])
    m = abi.dme_synth_motion(zoom, pan[0], pan[1], len(objects), objs, noise_sigma, noise_seed)
    out = np.empty_like(cur)
    check(lib().dme_synth_warp(ptr(cur), W, H, C.byref(m), ptr(out)))
    return out


def chroma(y, seed):
    y = np.ascontiguousarray(y, np.int16)
    H, W = y.shape
    u = np.empty((H // 2, W // 2), np.int16)
    v = np.empty((H // 2, W // 2), np.int16)
    check(lib().dme_synth_chroma(ptr(y), W, H, seed, ptr(u), ptr(v)))
    return u, v


def _objects(rng, W, H, n, dmax):
    objs = []
    for _ in range(n):
        w = W * rng.uniform(0.08, 0.3)
        h = H * rng.uniform(0.08, 0.3)
        x0 = rng.uniform(0, W - w)
        y0 = rng.uniform(0, H - h)
        objs.append((x0, y0, x0 + w, y0 + h, rng.uniform(-dmax, dmax), rng.uniform(-dmax, dmax)))
    return objs


def picture_set(seed, width, height, nrefs, pan_max, nobj, obj_max, zoom=1.0, noise=2.0):
    """Current picture (with chroma) and nrefs references.

    Reference 1 (the previous picture) is the current warped by motion M,
    reference 2 (the next picture) by the opposite motion.
    """
    rng = _rng(seed)
    cur = texture(seed * 7919 + 17, width, height, 3)
    u, v = chroma(cur, seed * 31 + 5)
    pan = (rng.uniform(-pan_max, pan_max), rng.uniform(-pan_max, pan_max))
    objs = _objects(rng, width, height, nobj, obj_max)
    refs = [warp(cur, zoom, pan, objs, noise, seed * 104729 + 1)]
    if nrefs > 1:
        neg = [(o[0], o[1], o[2], o[3], -o[4], -o[5]) for o in objs]
        refs.append(warp(cur, 2.0 - zoom, (-pan[0], -pan[1]), neg, noise, seed * 104729 + 2))
    return dict(cur=cur, u=u, v=v, refs=refs)


# Block parameters per picture size (xblen, yblen, xbsep, ybsep)
BLOCKS = {"cif": (12, 12, 8, 8), "720p": (12, 12, 8, 8), "1080p": (16, 16, 12, 12), "2160p": (24, 24, 16, 16)}
SIZES = {"cif": (352, 288), "720p": (1280, 720), "1080p": (1920, 1080), "2160p": (3840, 2160)}


def config0(seed):
    """CIF P picture, 1 ref, 12/8 blocks, 3-level pyramid, quarter pel, lambda 10."""
    d = picture_set(seed, 352, 288, 1, 6.0, 3, 12.0)
    d.update(block=BLOCKS["cif"], depth=3, precision=2, lam=10.0, nrefs=1, name="cif_p_1ref")
    return d


def config1(seed):
    """1080p B picture, 2 refs, 16/12 blocks, 4-level pyramid, quarter pel, lambda 10."""
    d = picture_set(seed, 1920, 1080, 2, 20.0, 8, 40.0, zoom=1.002)
    d.update(block=BLOCKS["1080p"], depth=4, precision=2, lam=10.0, nrefs=2, name="1080p_b_2ref")
    return d


def sized(seed, size, nrefs, depth=None):
    """One picture of a named size with the config-2 block parameters and
    motion scaled by resolution (config-1 motion at 1080p)."""
    W, H = SIZES[size]
    s = W / 1920.0
    d = picture_set(seed, W, H, nrefs, 20.0 * s, 8, 40.0 * s, zoom=1.002)
    dep = depth if depth is not None else {"cif": 3, "720p": 4, "1080p": 4, "2160p": 5}[size]
    d.update(block=BLOCKS[size], depth=dep, precision=2, lam=10.0, nrefs=nrefs, name="%s_%dref" % (size, nrefs))
    return d


def config2_mix(seed):
    """The 64-picture mixed batch: 16 CIF, 24 720p, 16 1080p, 8 2160p; 1 or
    2 refs by seed."""
    rng = _rng(seed)
    out = []
    for size, n in (("cif", 16), ("720p", 24), ("1080p", 16), ("2160p", 8)):
        for _ in range(n):
            out.append((size, int(rng.integers(1, 3)), int(rng.integers(0, 2 ** 31))))
    return out


def config4(seed):
    """2160p, 2 refs, pan up to +-128, 16 objects up to +-96, noise 3, 24/16 blocks, 5 levels."""
    d = picture_set(seed, 3840, 2160, 2, 128.0, 16, 96.0, zoom=1.002, noise=3.0)
    d.update(block=BLOCKS["2160p"], depth=5, precision=2, lam=10.0, nrefs=2, name="2160p_b_2ref")
    return d
