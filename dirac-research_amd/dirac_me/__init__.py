"""Python host mirror of the MI355X Dirac motion-estimation engine.

A thin layer over the C ABI of libdme.so (include/dme.h) used by the tests,
smoke() and bench.py.  ``Context`` mirrors the reference's libdirac_motionest
surface: ``pixel_search`` (PixelMatcher::DoSearch, pixel_match.h:81),
``set_lambda_map`` (MEData::SetLambdaMap, motion.h:362), ``subpel``
(SubpelRefine::DoSubpel, me_subpel.h:77), ``mode_decide``
(ModeDecider::DoModeDecn, me_mode_decn.h:93) and ``motion_estimate`` (the
live sequence of PictureCompressor, picture_compress.cpp:76-298).
"""
import ctypes as C
import os
import weakref

import numpy as np

from . import abi
from .abi import (COST_DTYPE, DME_FP_C, DME_FP_MMX, DME_INTRA, DME_REF1_ONLY, DME_REF2_ONLY,  # noqa: F401
                  DME_REF1AND2, MV_DTYPE, DmeError, check, lib, ptr)


def make_params(xblen=12, yblen=12, xbsep=8, ybsep=8, precision=2, depth=-1, lam=10.0, fp=DME_FP_MMX,
                full_search=False, x_range=32, y_range=32):
    p = abi.dme_params()
    p.xblen, p.yblen, p.xbsep, p.ybsep = xblen, yblen, xbsep, ybsep
    p.mv_precision = precision
    p.depth = depth
    p.lambda_ = lam
    p.fp_semantics = fp
    p.chroma_format = 0
    if full_search:
        p.full_search = 1
        p.x_range_me, p.y_range_me = x_range, y_range
    return p


def geometry(params, width, height):
    g = abi.dme_geometry()
    check(lib().dme_geometry_for(C.byref(params), width, height, C.byref(g)))
    return g


class MEData:
    """Host MEData arrays (motion.h:219-410) in the C-ABI layout."""

    FIELDS = ("vectors", "pred_costs", "mode", "intra_costs", "bipred_costs", "dc", "sb_split", "sb_costs",
              "lambda_map")

    def __init__(self, geom):
        Y, X = geom.ynum_blocks, geom.xnum_blocks
        YS, XS = geom.ynum_sb, geom.xnum_sb
        self.vectors = [np.zeros((Y, X), MV_DTYPE) for _ in range(2)]
        self.pred_costs = [np.zeros((Y, X), COST_DTYPE) for _ in range(2)]
        self.mode = np.zeros((Y, X), np.int32)
        self.intra_costs = np.zeros((Y, X), np.float32)
        self.bipred_costs = np.zeros((Y, X), COST_DTYPE)
        self.dc = [np.zeros((Y, X), np.int16) for _ in range(3)]
        self.sb_split = np.zeros((YS, XS), np.int32)
        self.sb_costs = np.zeros((YS, XS), np.float32)
        self.lambda_map = np.zeros((Y, X), np.float32)
        self.intra_block_ratio = np.zeros(1, np.float32)  # IntraModeAnalyse (picture_compress.cpp:300-320)

    def struct(self):
        m = abi.dme_medata()
        for r in range(2):
            m.vectors[r] = self.vectors[r].ctypes.data
            m.pred_costs[r] = self.pred_costs[r].ctypes.data
        m.mode = self.mode.ctypes.data
        m.intra_costs = self.intra_costs.ctypes.data
        m.bipred_costs = self.bipred_costs.ctypes.data
        for c in range(3):
            m.dc[c] = self.dc[c].ctypes.data
        m.sb_split = self.sb_split.ctypes.data
        m.sb_costs = self.sb_costs.ctypes.data
        m.lambda_map = self.lambda_map.ctypes.data
        m.intra_block_ratio = self.intra_block_ratio.ctypes.data
        return m

    def arrays(self, nrefs=2):
        """Flat dict of named arrays (for comparisons)."""
        d = {}
        for r in range(nrefs):
            d["vectors%d" % (r + 1)] = self.vectors[r]
            d["pred_costs%d" % (r + 1)] = self.pred_costs[r]
        d.update(mode=self.mode, intra_costs=self.intra_costs, bipred_costs=self.bipred_costs, dc_y=self.dc[0],
                 dc_u=self.dc[1], dc_v=self.dc[2], sb_split=self.sb_split, sb_costs=self.sb_costs,
                 lambda_map=self.lambda_map)
        return d


class Context:
    """A device context (dme_create / dme_destroy)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().dme_create(device, C.byref(h)))
        self.h = h
        self.params = None
        self._keep = {}
        self._graphs = weakref.WeakSet()

    def close(self):
        if self.h:
            for g in list(self._graphs):
                g.close()
            lib().dme_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        check(lib().dme_set_params(self.h, C.byref(params)))
        self.params = params

    def upload(self, pic_num, y, u=None, v=None):
        y = np.ascontiguousarray(y, np.int16)
        H, W = y.shape
        u = None if u is None else np.ascontiguousarray(u, np.int16)
        v = None if v is None else np.ascontiguousarray(v, np.int16)
        check(lib().dme_upload_picture(self.h, pic_num, W, H, ptr(y), ptr(u), ptr(v)))

    def bind_device(self, pic_num, width, height, d_y, d_u=None, d_v=None):
        check(lib().dme_bind_picture_device(self.h, pic_num, width, height, d_y, d_u, d_v))

    def drop(self, pic_num):
        check(lib().dme_drop_picture(self.h, pic_num))

    def _refs(self, refs):
        a = (C.c_int * 2)(*(list(refs) + [0] * (2 - len(refs))))
        return a, len(refs)

    def pixel_search(self, pic_num, refs, out=None):
        a, n = self._refs(refs)
        check(lib().dme_pixel_search(self.h, pic_num, a, n, C.byref(out.struct()) if out else None))
        return out

    def set_lambda_map(self, pic_num, lambda_map=None, out=None):
        lm = None if lambda_map is None else np.ascontiguousarray(lambda_map, np.float32)
        check(lib().dme_set_lambda_map(self.h, pic_num, ptr(lm), C.byref(out.struct()) if out else None))
        return out

    def load_medata(self, pic_num, refs, me, stage):
        """Hands the MEData state of `stage` (1 after the pel search, 2 with
        the lambda map, 3 after sub-pel refinement) back to the device."""
        r = (C.c_int * len(refs))(*refs)
        check(lib().dme_load_medata(self.h, pic_num, r, len(refs), C.byref(me.struct()), stage))

    def subpel(self, pic_num, out=None):
        check(lib().dme_subpel(self.h, pic_num, C.byref(out.struct()) if out else None))
        return out

    def mode_decide(self, pic_num, out=None):
        check(lib().dme_mode_decide(self.h, pic_num, C.byref(out.struct()) if out else None))
        return out

    def motion_estimate(self, pic_num, refs, out=None):
        a, n = self._refs(refs)
        check(lib().dme_motion_estimate(self.h, pic_num, a, n, C.byref(out.struct()) if out else None))
        return out

    @staticmethod
    def _jobs(jobs):
        arr = (abi.dme_job * len(jobs))()
        for i, (pn, refs, prm) in enumerate(jobs):
            arr[i].pic_num = pn
            arr[i].nrefs = len(refs)
            for k, r in enumerate(refs):
                arr[i].refs[k] = r
            arr[i].params = prm
        return arr

    def run_batch(self, jobs, outs=None):
        """jobs: list of (pic_num, refs, params)."""
        arr = self._jobs(jobs)
        if outs is not None:
            os_ = (abi.dme_medata * len(jobs))(*[o.struct() for o in outs])
            check(lib().dme_run_batch(self.h, arr, len(jobs), os_))
        else:
            check(lib().dme_run_batch(self.h, arr, len(jobs), None))
        return outs

    # ---- launch tuning and captured batch steps ----
    def set_tuning(self, parts=4, stagger=1, pel_wave_threads=1024, subpel_wave_threads=512, mode_wave_threads=1024):
        t = abi.dme_tuning(parts, stagger, pel_wave_threads, subpel_wave_threads, mode_wave_threads, 0.0)
        check(lib().dme_set_tuning(self.h, C.byref(t)))

    def reset_tuning(self):
        check(lib().dme_set_tuning(self.h, None))

    def get_tuning(self):
        t = abi.dme_tuning()
        check(lib().dme_get_tuning(self.h, C.byref(t)))
        return t

    def autotune_batch(self, jobs, reps=2):
        t = abi.dme_tuning()
        check(lib().dme_autotune_batch(self.h, self._jobs(jobs), len(jobs), reps, C.byref(t)))
        return t

    def save_tuning(self, path):
        check(lib().dme_tuning_save(self.h, os.fsencode(path)))

    def load_tuning(self, path):
        check(lib().dme_tuning_load(self.h, os.fsencode(path)))

    def batch_graph(self, jobs):
        """dme_batch_graph_create: a BatchGraph of these jobs."""
        return BatchGraph(self, jobs)

    def motion_estimate_noalloc(self, params, width, height, pic_num, d_planes, ref_nums, d_refs, out=None):
        """dme_motion_estimate_noalloc: d_planes = device addresses of the
        picture's Y, U, V; d_refs = device addresses of the references' Y;
        out = an MEData (host arrays) or a dme_medata struct (any memory)."""
        nums = (C.c_int * 2)(*(list(ref_nums) + [0] * (2 - len(ref_nums))))
        refs = (C.c_void_p * 2)(*(list(d_refs) + [None] * (2 - len(d_refs))))
        if isinstance(out, MEData):
            ms = out.struct()
        else:
            ms = out
        check(lib().dme_motion_estimate_noalloc(self.h, C.byref(params), width, height, pic_num,
                                                C.c_void_p(d_planes[0]), C.c_void_p(d_planes[1]),
                                                C.c_void_p(d_planes[2]), len(ref_nums), nums, refs,
                                                C.byref(ms) if ms is not None else None))
        return out

    def motion_compensate(self, params, width, height, me, refs, direction, planes):
        """MotionCompensator::CompensatePicture on host arrays: planes are
        three C-contiguous int16 arrays (Y, U, V; possibly padded), changed
        in place.  me: MEData with vectors, mode, dc and sb_split."""
        for a in planes:
            if a.dtype != np.int16 or not a.flags.c_contiguous:
                raise ValueError("planes must be C-contiguous int16 arrays")
        a, n = self._refs(refs)
        ps = (abi.dme_mc_plane * 3)(*[abi.dme_mc_plane(p.ctypes.data, p.shape[1], p.shape[0]) for p in planes])
        ms = me.struct()
        check(lib().dme_motion_compensate(self.h, C.byref(params), width, height, C.byref(ms), a, n, direction, ps))
        return planes

    def motion_compensate_batch(self, jobs):
        """jobs: list of dicts with pic_num, width, height, refs, direction,
        params, me (MEData or None: the context's estimate for pic_num) and
        planes: three (device pointer, width, height) triples."""
        arr = (abi.dme_mc_job * len(jobs))()
        keep = []
        for i, j in enumerate(jobs):
            arr[i].pic_num = j.get("pic_num", -1)
            arr[i].width, arr[i].height = j["width"], j["height"]
            arr[i].nrefs = len(j["refs"])
            for k, r in enumerate(j["refs"]):
                arr[i].refs[k] = r
            arr[i].direction = j["direction"]
            arr[i].params = j["params"]
            if j.get("me") is not None:
                ms = j["me"].struct()
                keep.append(ms)
                arr[i].mv = C.pointer(ms)
            for c, (d, w, h) in enumerate(j["planes"]):
                arr[i].planes[c] = abi.dme_mc_plane(d, w, h)
        check(lib().dme_motion_compensate_batch(self.h, arr, len(jobs)))

    def get_plane(self, pic_num, kind, level, shape):
        out = np.zeros(shape, np.int16)
        check(lib().dme_get_plane(self.h, pic_num, kind, level, ptr(out), out.size))
        return out

    def synchronize(self):
        check(lib().dme_synchronize(self.h))

    def stats(self):
        s = abi.dme_stats()
        check(lib().dme_get_stats(self.h, C.byref(s)))
        return s

    def kernel_times(self):
        """{kernel name: (ms, launches)} of the last call (profiling on)."""
        s = self.stats()
        out = {}
        for k in range(abi.DME_NUM_KERNELS):
            name = lib().dme_kernel_name(k).decode()
            if name and s.launches_by[k]:
                out[name] = (s.kernel_ms_by[k], s.launches_by[k])
        return out

    def set_profiling(self, on):
        check(lib().dme_set_profiling(self.h, 1 if on else 0))

    def invalidate(self, pic_num=-1):
        check(lib().dme_invalidate_caches(self.h, pic_num))

    # ---- learned proposals (include/dme_learned.h) ----
    def set_lp_weights(self, params):
        p = np.ascontiguousarray(params, np.float32)
        check(lib().dme_lp_set_weights(self.h, ptr(p), p.size))

    def motion_estimate_learned(self, pic_num, refs, out=None):
        a, n = self._refs(refs)
        check(lib().dme_motion_estimate_learned(self.h, pic_num, a, n, C.byref(out.struct()) if out else None))
        return out

    def run_batch_learned(self, jobs, outs=None):
        """jobs: list of (pic_num, refs, params)."""
        arr = (abi.dme_job * len(jobs))()
        for i, (pn, refs, prm) in enumerate(jobs):
            arr[i].pic_num = pn
            arr[i].nrefs = len(refs)
            for k, r in enumerate(refs):
                arr[i].refs[k] = r
            arr[i].params = prm
        os_ = None if outs is None else (abi.dme_medata * len(jobs))(*[o.struct() for o in outs])
        check(lib().dme_run_batch_learned(self.h, arr, len(jobs), os_))
        return outs

    def set_lp_form(self, form):
        check(lib().dme_lp_set_form(self.h, form))

    def lp_vectors(self, pic_num, which, geom, nrefs):
        """Level-0 integer vectors (which 0) or pre-merge sub-pel vectors
        (which 1) of a fast-form learned estimate, one array per reference."""
        out = [np.zeros((geom.ynum_blocks, geom.xnum_blocks), MV_DTYPE) for _ in range(nrefs)]
        p = [o.ctypes.data for o in out] + [None] * (2 - nrefs)
        check(lib().dme_lp_get_vectors(self.h, pic_num, which, C.c_void_p(p[0]), C.c_void_p(p[1])))
        return out

    def lp_outputs(self, pic_num, geom):
        """The network's 32 outputs per block of a fast-form learned estimate
        (valid until the context's next learned call)."""
        out = np.zeros((geom.ynum_blocks * geom.xnum_blocks, 32), np.float32)
        check(lib().dme_lp_get_outputs(self.h, pic_num, out.ctypes.data))
        return out

    def lp_stats(self):
        b, f = C.c_longlong(), C.c_longlong()
        check(lib().dme_lp_get_stats(self.h, C.byref(b), C.byref(f)))
        return b.value, f.value


class BatchGraph:
    """A captured batch step (dme_batch_graph_*)."""

    def __init__(self, ctx, jobs):
        self.ctx = ctx
        self.n = len(jobs)
        h = C.c_void_p()
        check(lib().dme_batch_graph_create(ctx.h, Context._jobs(jobs), len(jobs), C.byref(h)))
        self.h = h
        ctx._graphs.add(self)

    def launch(self, outs=None):
        if outs is not None:
            if len(outs) != self.n:
                raise ValueError("one MEData per job")
            os_ = (abi.dme_medata * self.n)(*[o.struct() for o in outs])
            check(lib().dme_batch_graph_launch(self.h, os_))
        else:
            check(lib().dme_batch_graph_launch(self.h, None))
        return outs

    def close(self):
        # dme_destroy destroys the graphs still alive: only destroy while
        # the context is open
        if self.h and self.ctx.h:
            check(lib().dme_batch_graph_destroy(self.h))
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lambda_map_host(vectors1, vectors2, lam, nrefs):
    v1 = np.ascontiguousarray(vectors1, MV_DTYPE)
    Y, X = v1.shape
    v2 = None if vectors2 is None else np.ascontiguousarray(vectors2, MV_DTYPE)
    out = np.zeros((Y, X), np.float32)
    check(lib().dme_lambda_map_host(X, Y, nrefs, ptr(v1), ptr(v2), lam, ptr(out)))
    return out
