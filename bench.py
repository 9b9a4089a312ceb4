"""Benchmark: plain GPU motion estimation of 1080p B pictures with two
references (16x16 blocks at separation 12, 4-level pyramid, quarter pel,
lambda 10), batched, on one MI355X.

A step is one pass of the whole live ME sequence (pyramids and up-converted
references rebuilt, pel search, lambda map, sub-pel refinement, mode
decision, DC) over one batch of synthetic pictures resident in HBM.  The
headline replays that step captured as a HIP graph (dme_batch_graph_launch)
under the launch setting dme_autotune_batch picks; the eager dme_run_batch
step is reported beside it ("eager", "tuned_and_graph").  Prints
one JSON line; the compulsory bytes and algorithmic SAD terms behind the
roofline figures come from _pel_bulk_bytes() and _sad_terms().  The line also carries the learned-proposal
leg, the motion-compensation leg (residuals of the same batch from its
estimates, mc_leg()), the full-search leg (full_search_leg()), the .imt
diagnostics-text leg (imt_leg()), the mixed-size batch (mixed_leg()) and the
8 x 2160p large-motion batch, plain against learned (large_motion_leg()).
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dirac-research_amd"))

METRIC = "1080p 2-ref ME frames/s: plain GPU vs learned-proposal at ≤1% mean cost gap"
HBM_PEAK_GBS = 8000.0
SAD_TERMS_PEAK = 256 * 64 * 2.4e9 * 4  # v_sad_u8: 4 absolute differences per lane-op


def _pel_bulk_bytes(geom, nrefs):
    """Compulsory HBM bytes of k_pel_bulk per picture, summed over the levels
    (one launch per level): every sample of the level's picture and of its
    references read once (level 0: the byte planes, 1 byte a sample; the
    pyramid levels: int16), each block's guide vector read (8 bytes) and its
    28-int pel record written, per reference."""
    total = 0.0
    for l in range(geom.depth + 1):
        n = geom.level_width[l] * geom.level_height[l]
        nb = geom.level_xblocks[l] * geom.level_yblocks[l]
        total += (1 + nrefs) * n * (1 if l == 0 else 2) + nrefs * nb * (28 * 4 + (8 if l < geom.depth else 0))
    return total


def _sad_terms():
    """Per-stage algorithmic SAD terms (N_sad) per picture of the benchmark
    pictures, counted by the oracle when the fixture was made
    (tests/golden/nsad_1080p.json, tests/golden/make_nsad.py): mean over
    the 8 pictures of the batch."""
    with open(os.path.join(ROOT, "tests", "golden", "nsad_1080p.json")) as f:
        rows = json.load(f)["frames"]
    return {k: sum(r[k] for r in rows) / len(rows) for k in ("pel", "subpel", "mode")}


def _profiled_traffic(kernel, batch):
    """HBM bytes per launch of `kernel` measured with rocprofv3 counters on
    this workload, from the summary kept in profiles/ (None if absent)."""
    for name in ("pel_bulk_1080p_2ref_batch64.md", "mc_1080p_2ref_batch64.md"):
        try:
            with open(os.path.join(ROOT, "profiles", name)) as f:
                for line in f:
                    if line.startswith("traffic_bytes_per_launch:"):
                        _, rest = line.split(":", 1)
                        fields = dict(kv.strip().split("=") for kv in rest.split(","))
                        if fields.get("kernel") == kernel and int(fields.get("batch", -1)) == batch:
                            return float(fields["bytes"])
        except (OSError, ValueError, KeyError):
            continue
    return None


def _copy_calibration(mb):
    """Device-to-device copy of `mb` MiB (the HIP runtime's copy kernel), so
    that a counter run reads FETCH_SIZE / WRITE_SIZE of a kernel whose traffic
    is known."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mb << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def _cpu_model():
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return cpu


def _cpu_threads():
    """Host threads for the N-core CPU baseline: the cores this process may
    use, at most 16 (the host share of one GPU on the box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(frames, params, budget_s):
    """Reference encoder ME (oracle/_ref, compiled from the reference's
    sources) or the oracle, on a bounded sample of the same workload: on one
    thread, and on N threads with one picture per thread (the reference is
    single-threaded per picture; pictures are independent).  The top-level
    figure is the N-thread one."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O

    kind = "reference" if O.ref_available() else "port"
    run = O.run_ref if kind == "reference" else O.run_oracle
    what = "reference encoder ME (MMX build, compiled from its sources)" if kind == "reference" else "oracle port"

    def one(i):
        d = frames[i % len(frames)]
        run(params, d["cur"], d["refs"], d["u"], d["v"])

    n = 0
    t0 = time.time()
    while True:
        one(n)
        n += 1
        if time.time() - t0 >= budget_s / 2 or n >= 200:
            break
    dt1 = time.time() - t0
    T = _cpu_threads()
    m = 0
    t0 = time.time()
    with ThreadPoolExecutor(max_workers=T) as ex:
        while True:
            list(ex.map(one, range(m, m + T)))
            m += T
            if time.time() - t0 >= budget_s / 2 or m >= 50 * T:
                break
    dtn = time.time() - t0
    cpu = _cpu_model()
    return {"value": m / dtn, "unit": "frames/s", "cores": T, "kind": kind,
            "sample": "%d 1080p B pictures (2 refs) of the same synthetic workload, %s, %d threads (one picture "
                      "per thread), %.1f s on %s" % (m, what, T, dtn, cpu),
            "single_core": {"value": n / dt1, "unit": "frames/s", "cores": 1,
                            "sample": "%d pictures, one thread, %.1f s" % (n, dt1)}}


def _mc_bytes(geom, me):
    """Algorithmic HBM bytes of compensating one picture (k_mc), counted per
    unit with no reuse between units, as the reference computes it
    (CompensateBlock per unit, mot_comp.cpp:541-603): each unit's OBMC
    window (clipped to the picture) reads 4 up-converted samples per sample,
    one byte each (the planes hold samples + 128), from each reference its
    mode uses (none for DC, two for bi-prediction); every sample of the three
    arrays is read and written once (int16); plus the MvData."""
    import numpy as np
    total = 0.0
    xnb, ynb, xsb = geom.xnum_blocks, geom.ynum_blocks, geom.xnum_sb
    refs_of = np.array([0, 1, 1, 2, 0, 0, 0, 0])
    for c in range(2):
        xbl, ybl = (geom.luma_blen[2] if c == 0 else geom.chroma_blen[2])
        xbs, ybs = (geom.luma_bsep[2] if c == 0 else geom.chroma_bsep[2])
        pw, ph = (geom.width, geom.height) if c == 0 else (geom.chroma_width, geom.chroma_height)
        xoff, yoff, xov = (xbl - xbs) // 2, (ybl - ybs) // 2, xbl - xbs
        nbx = min((pw + xbs + xoff - 1) // xbs, xnb)
        nby = min((ph + ybs + yoff - 1) // ybs, ynb)
        area_refs = 0
        for rb in range(nby):
            py = rb * ybs - yoff
            rows = min(py + ybl, ph) - max(py, 0)
            u = 0
            while u < nbx:
                sp = int(me.sb_split[rb // 4, u // 4])
                k = 4 >> (sp if sp in (0, 1) else 2)
                px = u * xbs - xoff
                cols = min(px + k * xbs + xov, pw) - max(px, 0)
                area_refs += rows * cols * refs_of[int(me.mode[rb, u]) & 7]
                u += k
        total += area_refs * 4 * (1 if c == 0 else 2) + pw * ph * 4 * (1 if c == 0 else 2)
    B = xnb * ynb
    return total + B * (2 * 8 + 4 + 3 * 2) + xsb * geom.ynum_sb * 4


def export_leg(ctx, jobs, geom, B, args, step_s):
    """The same step with every picture's MEData copied to host memory (the
    arrays a drop-in caller reads; about 1 MB per 1080p picture), so that
    the cost the headline leaves on the device is visible beside it."""
    import dirac_me

    outs = [dirac_me.MEData(geom) for _ in range(B)]

    def step():
        ctx.invalidate(-1)
        ctx.run_batch(jobs, outs)

    step()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    nbytes = sum(a.nbytes for a in outs[0].arrays(2).values())
    return {"value": B / dt, "unit": "frames/s", "ms_per_step": dt * 1e3,
            "export_ms_per_step": (dt - step_s) * 1e3, "medata_bytes_per_picture": nbytes,
            "note": "PCIe-inclusive: MEData of every picture copied to pageable host arrays each step"}


def mc_leg(ctx, jobs, frames, params, geom, B, args):
    """Motion compensation (subtract: the encoder's residual) of the batch's
    pictures from their device-side estimates, all pictures in one
    dme_motion_compensate_batch call per step; the step rebuilds the
    references' up-converted planes first (the reference builds them once
    per reference picture, Picture::UpData).  Also the reference's own
    MotionCompensator on the host, on a bounded sample."""
    import numpy as np
    import dirac_me
    from dirac_me import abi
    from dirac_me.learned import DeviceArray

    est = [dirac_me.MEData(geom) for _ in frames]
    ctx.run_batch(jobs[:len(frames)], est)
    arrays, mjobs = [], []
    for j, (pn, refs, prm) in enumerate(jobs):
        d = frames[j % len(frames)]
        src = (d["cur"], d["u"], d["v"])
        dp = []
        for a in src:
            x = DeviceArray(ctx, a.shape, np.int16)
            x.upload(a)
            dp.append(x)
        arrays.append(dp)
        mjobs.append(dict(pic_num=pn, width=1920, height=1080, refs=refs, direction=abi.DME_MC_SUBTRACT,
                          params=prm, me=None, planes=[(x.ptr, a.shape[1], a.shape[0]) for x, a in zip(dp, src)]))

    def step():
        ctx.invalidate(-1)
        ctx.motion_compensate_batch(mjobs)

    for _ in range(args.warmup):
        step()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    dt = (time.perf_counter() - t0) / args.steps
    ctx.set_profiling(True)
    step()
    kt = ctx.kernel_times()
    ctx.set_profiling(False)
    ms, nl = kt["k_mc"]
    alg = sum(_mc_bytes(geom, est[j % len(frames)]) for j in range(B))
    achieved = alg / nl / (ms / nl * 1e-3) / 1e9
    out = {"value": B / dt, "unit": "pictures/s", "ms_per_step": dt * 1e3, "batch": B,
           "direction": "subtract", "mv_source": "device-side estimates of the same batch",
           "roofline": {"kernel": "k_mc", "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": _profiled_traffic("k_mc", B),
                        "algorithmic_bytes_per_launch": alg / nl, "avg_launch_ms": ms / nl},
           "kernel_ms_per_step": {k: round(v[0], 4) for k, v in kt.items()}}
    for dp in arrays:
        for x in dp:
            x.free()
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        kind = "reference" if O.ref_available() else "port"
        n = 0
        t0 = time.time()
        while True:
            d = frames[n % len(frames)]
            planes = [d["cur"].copy(), d["u"].copy(), d["v"].copy()]
            O.motion_compensate("mmx" if kind == "reference" else "oracle", params, d["refs_yuv"],
                                est[n % len(frames)], 2, abi.DME_MC_SUBTRACT, planes)
            n += 1
            if time.time() - t0 >= args.cpu_budget / 2 or n >= 100:
                break
        cdt = time.time() - t0
        out["cpu_baseline"] = {"value": n / cdt, "unit": "pictures/s", "cores": 1, "kind": kind,
                               "sample": "%d 1080p B pictures (2 refs), %s, single thread, %.1f s, references' "
                                         "up-conversion included" % (
                                             n, "reference MotionCompensator (MMX build, compiled from its sources)"
                                             if kind == "reference" else "oracle port", cdt)}
    return out


def imt_leg(ctx, jobs, frames, geom, B, args):
    """The .imt diagnostics text of the batch's estimates (host code, one
    thread): dme_imt_write_picture, then dme_imt_read_picture, over the B
    1080p two-reference records of one step.  Also the reference's own
    operator<<(ostream&, MEData&) on a bounded sample."""
    import dirac_me
    from dirac_me import imt

    est = [dirac_me.MEData(geom) for _ in frames]
    ctx.run_batch(jobs[:len(frames)], est)
    t0 = time.perf_counter()
    recs = [imt.write_picture(jobs[j][0], geom, jobs[j][1], est[j % len(frames)]) for j in range(B)]
    tw = time.perf_counter() - t0
    nbytes = sum(len(r) for r in recs)
    t0 = time.perf_counter()
    for r in recs:
        imt.read_picture(r)
    tr = time.perf_counter() - t0
    out = {"write": {"value": B / tw, "unit": "records/s", "MB_per_s": nbytes / tw / 1e6},
           "read": {"value": B / tr, "unit": "records/s", "MB_per_s": nbytes / tr / 1e6},
           "bytes_per_record": nbytes / B, "records": B, "threads": 1}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        kind = "reference" if O.ref_available() else "port"
        n = 0
        t0 = time.time()
        while True:
            me = est[n % len(frames)]
            if kind == "reference":
                O.ref_imt_medata_text(me, 2)
            else:
                O.imt_record(jobs[n][0], jobs[n][1], geom.luma_bsep[2][1], geom.luma_bsep[2][0], me)
            n += 1
            if time.time() - t0 >= args.cpu_budget / 4 or n >= B:
                break
        cdt = time.time() - t0
        out["cpu_baseline"] = {"value": n / cdt, "unit": "records/s", "cores": 1, "kind": kind,
                               "sample": "%d 1080p 2-ref records, %s, %.1f s" % (
                                   n, "the reference's operator<< on MEData (compiled from its sources), "
                                   "MEData filled from the arrays included" if kind == "reference"
                                   else "oracle port (Python)", cdt)}
    return out


def mvdata_leg(ctx, jobs, frames, geom, B, args):
    """The MV-data wire format of the batch's estimates (host code):
    dme_mvdata_encode_batch over the B 1080p two-reference pictures of one
    step on one thread and on 16, and dme_mvdata_decode on one thread.  The
    reference's coder cannot be built here (it needs the configure-generated
    dirac-stdint.h), so the oracle's pure-Python restatement is timed beside
    it on a bounded sample."""
    import dirac_me
    from dirac_me import mvdata

    est = [dirac_me.MEData(geom) for _ in frames]
    ctx.run_batch(jobs[:len(frames)], est)
    infos = [mvdata.info(geom, 2, 2) for _ in range(B)]
    mes = [est[j % len(frames)] for j in range(B)]
    mvdata.encode_batch(infos[:2], mes[:2], threads=1, cap=1 << 20)
    out = {}
    for th in (1, 16):
        t0 = time.perf_counter()
        data = mvdata.encode_batch(infos, mes, threads=th, cap=1 << 20)
        dt = time.perf_counter() - t0
        out["encode_%d_threads" % th] = {"value": B / dt, "unit": "pictures/s", "threads": th}
    nbytes = sum(len(x) for x in data)
    t0 = time.perf_counter()
    for x in data:
        mvdata.decode(x, geom, 2)
    dt = time.perf_counter() - t0
    out["decode"] = {"value": B / dt, "unit": "pictures/s", "threads": 1}
    out["bytes_per_picture"] = nbytes / B
    out["pictures"] = B
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import mvcodec_ref as R
        me = est[0]
        split = me.sb_split.tolist()
        mode = me.mode.tolist()
        vec = [[[(int(v["x"]), int(v["y"])) for v in row] for row in me.vectors[r]] for r in range(2)]
        dc = [me.dc[c].tolist() for c in range(3)]
        bp = tuple(geom.luma_blen[2]) + tuple(geom.luma_bsep[2])
        n = 0
        t0 = time.time()
        while n < 4 and time.time() - t0 < args.cpu_budget / 4:
            R.encode(split, mode, vec, dc, 2, bp, 2)
            n += 1
        cdt = time.time() - t0
        out["cpu_baseline"] = {"value": n / cdt, "unit": "pictures/s", "cores": 1, "kind": "port",
                               "sample": "%d 1080p 2-ref pictures, the oracle's pure-Python restatement (the "
                                         "reference's coder cannot be built here), %.1f s" % (n, cdt)}
    return out


def full_search_leg(ctx, jobs, frames, args):
    """Full search (the encoder's FullSearch option, range 32: a 65 x 65
    window per block and reference at temporal distance 1) over the first
    16 pictures of the batch, the whole live sequence per step; the
    reference's own full search on the host beside it."""
    import dirac_me

    n = min(16, len(jobs))
    p = dirac_me.make_params(*frames[0]["block"], precision=2, lam=10.0, full_search=True, x_range=32, y_range=32)
    fj = [(pn, refs, p) for pn, refs, _ in jobs[:n]]
    for _ in range(max(1, args.warmup)):
        ctx.invalidate(-1)
        ctx.run_batch(fj)
    ctx.synchronize()
    steps = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.invalidate(-1)
        ctx.run_batch(fj)
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / steps
    ctx.set_profiling(True)
    ctx.invalidate(-1)
    ctx.run_batch(fj)
    kt = ctx.kernel_times()
    ctx.set_profiling(False)
    g = dirac_me.geometry(p, 1920, 1080)
    xl, yl = g.luma_blen[2]
    # SAD terms of k_pel_full: every block and reference, the 65 x 65 window
    # plus the 5 x 5 neighbourhood of zero, block samples clipped to the picture
    terms = 0
    for yb in range(g.ynum_blocks):
        for xb in range(g.xnum_blocks):
            x0, y0 = xb * g.luma_bsep[2][0] - (xl - g.luma_bsep[2][0]) // 2, yb * g.luma_bsep[2][1] - (yl - g.luma_bsep[2][1]) // 2
            w = min(x0 + xl, 1920) - max(x0, 0)
            h = min(y0 + yl, 1080) - max(y0, 0)
            if w > 0 and h > 0:
                terms += w * h
    terms *= (65 * 65 + 25) * 2 * n
    ms = kt["k_pel_full"][0]
    out = {"value": n / dt, "unit": "frames/s", "ms_per_step": dt * 1e3, "batch": n, "range": 32,
           "k_pel_full_ms": ms, "sad_terms_per_s": terms / (ms * 1e-3),
           "sad_frac": terms / (ms * 1e-3) / SAD_TERMS_PEAK,
           "kernel_ms_per_step": {k: round(v[0], 4) for k, v in kt.items()}}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        kind = "reference" if O.ref_available() else "port"
        run = O.run_ref if kind == "reference" else O.run_oracle
        d = frames[0]
        t0 = time.time()
        run(p, d["cur"], d["refs"], d["u"], d["v"], pic_nums=(1, [0, 2]))
        cdt = time.time() - t0
        out["cpu_baseline"] = {"value": 1.0 / cdt, "unit": "frames/s", "cores": 1, "kind": kind,
                               "sample": "1 1080p B picture (2 refs, range 32), %s, single thread, %.1f s" % (
                                   "reference encoder ME with FullSearch (MMX build)" if kind == "reference"
                                   else "oracle port", cdt)}
    return out


def learned_leg(ctx, jobs, frames, geom, B, args, plain_fps):
    """The same batch through learned proposals (dme_run_batch_learned):
    frames/s, and the Lagrangian cost gap against the plain search (sum of
    SBCosts, learned / plain - 1) over the distinct pictures of the batch."""
    import numpy as np
    import dirac_me
    from dirac_me import learned

    ctx.set_lp_weights(learned.load_weights())
    for _ in range(args.warmup):
        ctx.invalidate(-1)
        ctx.run_batch_learned(jobs)
    ctx.synchronize()
    b0, f0 = ctx.lp_stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.invalidate(-1)
        ctx.run_batch_learned(jobs)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    b1, f1 = ctx.lp_stats()
    n = len(frames)
    plain = [dirac_me.MEData(geom) for _ in range(n)]
    lrn = [dirac_me.MEData(geom) for _ in range(n)]
    ctx.invalidate(-1)
    ctx.run_batch(jobs[:n], plain)
    ctx.invalidate(-1)
    ctx.run_batch_learned(jobs[:n], lrn)
    gaps = [float(np.sum(l.sb_costs, dtype=np.float64) / np.sum(p.sb_costs, dtype=np.float64) - 1.0)
            for p, l in zip(plain, lrn)]
    fps = B * args.steps / dt
    ctx.set_profiling(True)
    ctx.invalidate(-1)
    ctx.run_batch_learned(jobs)
    ctx.synchronize()
    kt = ctx.kernel_times()
    ctx.set_profiling(False)
    return {"value": fps, "unit": "frames/s", "ms_per_step": dt / args.steps * 1e3, "form": "fast",
            "plain_for_speedup": "eager dme_run_batch, default setting",
            "kernel_ms_per_step": {k: round(v[0], 4) for k, v in kt.items()},
            "speedup_vs_plain": fps / plain_fps, "mean_sb_cost_gap": float(np.mean(gaps)),
            "max_sb_cost_gap": float(np.max(gaps)), "fallback_rate": (f1 - f0) / max(1, b1 - b0),
            "pictures_for_gap": n, "weights": "dirac_me/proposal_weights.npy (2000 training pictures)"}


def combined_leg(ctx, jobs, args):
    """The same batch with combined-component ME (EncoderParams::CombinedME:
    every stage on the combined luma+chroma pictures, 16-bit samples):
    frames/s of the eager step."""
    import copy

    cjobs = []
    for pn, refs, prm in jobs:
        q = copy.copy(prm)
        q.combined_me = 1
        cjobs.append((pn, refs, q))

    def step():
        ctx.invalidate(-1)
        ctx.run_batch(cjobs)

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    ctx.set_profiling(True)
    step()
    ctx.synchronize()
    kt = ctx.kernel_times()
    ctx.set_profiling(False)
    return {"value": len(cjobs) / dt, "unit": "frames/s", "ms_per_step": dt * 1e3,
            "kernel_ms_per_step": {k: round(v[0], 4) for k, v in kt.items()}}


def tuned_and_graph(ctx, jobs, args, eager_ms):
    """The same batch step with the setting dme_autotune_batch picks (eager
    dme_run_batch), and replayed as a captured HIP graph of that setting
    (dme_batch_graph_launch): ms per step of both.  Every form does the whole
    step: planes rebuilt from the pictures' luma, all stages, the host's
    lambda check."""
    t = ctx.autotune_batch(jobs, reps=2)

    def timed(step):
        for _ in range(args.warmup):
            step()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        ctx.synchronize()
        return (time.perf_counter() - t0) / args.steps * 1e3

    def eager():
        ctx.invalidate(-1)
        ctx.run_batch(jobs)

    tuned_ms = timed(eager)
    g = ctx.batch_graph(jobs)
    try:
        graph_ms = timed(g.launch)
    finally:
        g.close()
    setting = {"parts": t.parts, "stagger": t.stagger, "pel_wave_threads": t.pel_wave_threads,
               "subpel_wave_threads": t.subpel_wave_threads, "mode_wave_threads": t.mode_wave_threads}
    return {"eager_default_ms": eager_ms, "tuning": setting, "eager_tuned_ms": tuned_ms, "graph_tuned_ms": graph_ms,
            "frames_per_s": {"eager_default": len(jobs) / eager_ms * 1e3, "eager_tuned": len(jobs) / tuned_ms * 1e3,
                             "graph_tuned": len(jobs) / graph_ms * 1e3}}


def mixed_leg(args):
    """The mixed batch of 64 pictures (16 CIF, 24 720p, 16 1080p, 8 2160p,
    1 or 2 references by seed, block parameters per size) in one
    dme_run_batch call per step: frames/s and luma samples/s; then autotuned
    and graph-captured (tuned_and_graph)."""
    from concurrent.futures import ThreadPoolExecutor
    import dirac_me
    from dirac_me import synth

    mix = synth.config2_mix(0)
    with ThreadPoolExecutor(max_workers=8) as ex:
        sets = list(ex.map(lambda m: synth.sized(m[2], m[0], m[1]), mix))
    ctx = dirac_me.Context(0)
    jobs = []
    samples = 0
    for j, ((size, nrefs, seed), d) in enumerate(zip(mix, sets)):
        prm = dirac_me.make_params(*d["block"], precision=2, depth=d["depth"], lam=10.0)
        base = 3 * j
        nums = [base, base + 2][:nrefs]
        for n, r in zip(nums, d["refs"]):
            ctx.upload(n, r)
        ctx.upload(base + 1, d["cur"], d["u"], d["v"])
        jobs.append((base + 1, nums, prm))
        samples += d["cur"].size
    del sets

    def step():
        ctx.invalidate(-1)
        ctx.run_batch(jobs)

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    ctx.set_profiling(True)
    step()
    ctx.synchronize()
    kt = ctx.kernel_times()
    ctx.set_profiling(False)
    forms = tuned_and_graph(ctx, jobs, args, dt * 1e3)
    ctx.close()
    return {"metric": "mixed-size ME frames/s (64 pictures per batch)", "value": len(jobs) / dt, "unit": "frames/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int16", "data": "synthetic",
            "config": {"workload": "mixed batch: 16 CIF, 24 1280x720, 16 1920x1080, 8 3840x2160, 1 or 2 refs by seed, "
                                   "blocks 12/8, 12/8, 16/12, 24/16, quarter pel, lambda 10", "batch": len(jobs)},
            "luma_samples_per_s": samples / dt, "tuned_and_graph": forms,
            "kernel_ms_per_step": {k: round(v[0], 4) for k, v in kt.items()}}


def large_motion_leg(args):
    """Large-motion stress: 8 pictures of 3840 x 2160 per batch (2 refs, pan
    up to +-128, 16 objects up to +-96, noise 3, 24/16 blocks, 5 levels,
    quarter pel), plain batch against the learned batch: frames/s of both,
    the speedup and the SBCosts gap over the 8 pictures."""
    from concurrent.futures import ThreadPoolExecutor
    import numpy as np
    import dirac_me
    from dirac_me import learned, synth

    with ThreadPoolExecutor(max_workers=8) as ex:
        sets = list(ex.map(synth.config4, range(700, 708)))
    ctx = dirac_me.Context(0)
    ctx.set_lp_weights(learned.load_weights())
    prm = dirac_me.make_params(*sets[0]["block"], precision=2, depth=5, lam=10.0)
    geom = dirac_me.geometry(prm, 3840, 2160)
    jobs = []
    for j, d in enumerate(sets):
        base = 3 * j
        ctx.upload(base, d["refs"][0])
        ctx.upload(base + 2, d["refs"][1])
        ctx.upload(base + 1, d["cur"], d["u"], d["v"])
        jobs.append((base + 1, [base, base + 2], prm))
    del sets
    out = {}
    for form, run in (("plain", ctx.run_batch), ("learned", ctx.run_batch_learned)):
        for _ in range(args.warmup):
            ctx.invalidate(-1)
            run(jobs)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.invalidate(-1)
            run(jobs)
        ctx.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        ctx.set_profiling(True)
        ctx.invalidate(-1)
        run(jobs)
        ctx.synchronize()
        kt = ctx.kernel_times()
        ctx.set_profiling(False)
        out[form] = {"value": len(jobs) / dt, "unit": "frames/s", "ms_per_step": dt * 1e3,
                     "kernel_ms_per_step": {k: round(v[0], 4) for k, v in kt.items()}}
    res = {f: [dirac_me.MEData(geom) for _ in jobs] for f in ("plain", "learned")}
    ctx.invalidate(-1)
    ctx.run_batch(jobs, res["plain"])
    b0, f0 = ctx.lp_stats()
    ctx.invalidate(-1)
    ctx.run_batch_learned(jobs, res["learned"])
    b1, f1 = ctx.lp_stats()
    ctx.close()
    gaps = [float(np.sum(l.sb_costs, dtype=np.float64) / np.sum(p.sb_costs, dtype=np.float64) - 1.0)
            for p, l in zip(res["plain"], res["learned"])]
    return {"config": {"workload": "8 x 3840x2160 B pictures per batch, 2 refs, pan up to +-128, 16 objects up to "
                                   "+-96, noise 3, blocks 24/16, 5 levels, quarter pel, lambda 10", "batch": len(jobs)},
            "plain": out["plain"], "learned": out["learned"],
            "speedup_vs_plain": out["learned"]["value"] / out["plain"]["value"],
            "mean_sb_cost_gap": float(np.mean(gaps)), "max_sb_cost_gap": float(np.max(gaps)),
            "fallback_rate": (f1 - f0) / max(1, b1 - b0)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-mc", action="store_true")
    ap.add_argument("--no-full", action="store_true")
    ap.add_argument("--no-imt", action="store_true")
    ap.add_argument("--no-mvdata", action="store_true", help="skip the MV-data wire-format leg")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-large", action="store_true", help="skip the 8 x 2160p large-motion leg")
    ap.add_argument("--no-combined", action="store_true", help="skip the combined-component leg")
    ap.add_argument("--learned-only", action="store_true",
                    help="only the learned-proposal leg of the 1080p workload (for per-kernel profiles)")
    ap.add_argument("--workload", default="1080p", choices=("1080p", "mixed64"),
                    help="1080p: batches of 1080p B pictures (default); mixed64: the mixed-size batch")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run a device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the motion-estimation path runs one picture batch per GPU and "
              "there is no multi-GPU launcher (pictures would shard by picture with no exchange)",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "mixed64":
        print(json.dumps(mixed_leg(args)))
        return

    import dirac_me
    from dirac_me import synth

    B = args.batch
    frames = [synth.config1(100 + s) for s in range(min(B, 8))]  # seeds outside the proposal training set
    d0 = frames[0]
    params = dirac_me.make_params(*d0["block"], precision=2, depth=4, lam=10.0)
    geom = dirac_me.geometry(params, 1920, 1080)
    ctx = dirac_me.Context(0)
    ctx.set_params(params)
    jobs = []
    for i, d in enumerate(frames):
        d["refs_yuv"] = [(r,) + synth.chroma(r, 500 + 2 * i + k) for k, r in enumerate(d["refs"])]
    for j in range(B):
        d = frames[j % len(frames)]
        base = 3 * j
        ctx.upload(base, *d["refs_yuv"][0])  # reference chroma: used by motion compensation only
        ctx.upload(base + 2, *d["refs_yuv"][1])
        ctx.upload(base + 1, d["cur"], d["u"], d["v"])
        jobs.append((base + 1, [base, base + 2], params))
    print("bench: %d pictures uploaded; warmup %d, steps %d" % (B, args.warmup, args.steps), file=sys.stderr, flush=True)

    def step():
        ctx.invalidate(-1)
        ctx.run_batch(jobs)

    if args.learned_only:
        print(json.dumps({"learned": learned_leg(ctx, jobs, frames, geom, B, args, float("nan"))}))
        ctx.close()
        return
    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.synchronize()
    dt = time.perf_counter() - t0
    fps = B * args.steps / dt
    print("bench: timed %.3f s" % dt, file=sys.stderr, flush=True)

    # one more step with per-kernel events, for the roofline figures
    ctx.set_profiling(True)
    step()
    ctx.synchronize()
    kt = ctx.kernel_times()
    ctx.set_profiling(False)
    # the dominant kernel: compulsory HBM bytes per launch against its mean
    # launch time (HBM basis), and the pel stage's algorithmic SAD terms
    # against the v_sad_u8 peak (the integer-VALU basis that bounds it)
    dom = max(kt, key=lambda k: kt[k][0])
    ms, nl = kt[dom]
    per_launch_ms = ms / nl
    alg_bytes = _pel_bulk_bytes(geom, 2) * B if dom == "k_pel_bulk" else None
    roof = {"kernel": dom, "bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "avg_launch_ms": per_launch_ms,
            "launches_per_step": nl, "traffic": _profiled_traffic(dom, B)}
    if alg_bytes is not None:
        achieved = alg_bytes / nl / (per_launch_ms * 1e-3) / 1e9
        roof.update(achieved=achieved, frac=achieved / HBM_PEAK_GBS, algorithmic_bytes_per_launch=alg_bytes / nl)
    nsad = _sad_terms()
    stages = {"pel": ("k_pel_bulk", "k_pel_wave"), "subpel": ("k_subpel_bulk", "k_subpel_wave"),
              "mode": ("k_mode_l2", "k_mode_tables", "k_mode_wave")}
    sad = {}
    for st, ks in stages.items():
        t = sum(kt[k][0] for k in ks if k in kt) * 1e-3
        if t > 0:
            a = nsad[st] * B / t
            sad[st] = {"kernels": list(ks), "sad_terms_per_step": nsad[st] * B, "achieved": a,
                       "frac": a / SAD_TERMS_PEAK}
    a = nsad["pel"] * B / (kt["k_pel_bulk"][0] * 1e-3)
    roof_sad = {"kernel": "k_pel_bulk", "bound": "valu", "achieved": a, "peak": SAD_TERMS_PEAK,
                "unit": "SAD terms/s", "frac": a / SAD_TERMS_PEAK,
                "terms_per_launch": nsad["pel"] * B / kt["k_pel_bulk"][1],
                "note": "the pel stage's algorithmic SAD terms (N_sad, counted by the oracle) over k_pel_bulk's time; "
                        "peak = 256 CUs x 64 lanes x 2.4 GHz x 4 terms per v_sad_u8"}
    # the headline form: the captured batch step with the autotuned setting
    # (the same whole step; the eager dme_run_batch figures stand beside it)
    forms = tuned_and_graph(ctx, jobs, args, dt / args.steps * 1e3)
    ctx.reset_tuning()
    gfps = forms["frames_per_s"]["graph_tuned"]
    out = {
        "metric": METRIC, "value": gfps, "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": forms["graph_tuned_ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int16", "data": "synthetic",
        "config": {"workload": "plain GPU ME, 1920x1080 4:2:0 B pictures with 2 refs, blocks 16/12, 4-level "
                               "pyramid, quarter pel, lambda 10; batch of %d pictures per step" % B,
                   "batch": B, "width": 1920, "height": 1080, "refs": 2, "depth": 4, "precision": "quarter",
                   "step": "captured batch step (dme_batch_graph_launch), autotuned launch setting"},
        "block_decisions_per_s": gfps * geom.xnum_blocks * geom.ynum_blocks * 2,
        "eager": {"value": fps, "ms_per_step": dt / args.steps * 1e3, "step": "dme_run_batch, default setting"},
        "tuned_and_graph": forms,
        "roofline": roof, "roofline_sad": roof_sad, "sad_terms_by_stage": sad,
        "kernel_ms_per_step": {k: round(v[0], 4) for k, v in kt.items()},
    }
    out["with_export"] = export_leg(ctx, jobs, geom, B, args, dt / args.steps)
    if not args.no_combined:
        out["combined_me"] = combined_leg(ctx, jobs, args)
    if not args.no_mc:
        out["motion_compensation"] = mc_leg(ctx, jobs, frames, params, geom, B, args)
    if not args.no_full:
        out["full_search"] = full_search_leg(ctx, jobs, frames, args)
    if not args.no_imt:
        out["imt"] = imt_leg(ctx, jobs, frames, geom, B, args)
    if not args.no_mvdata:
        out["mvdata"] = mvdata_leg(ctx, jobs, frames, geom, B, args)
    if not args.no_learned:
        out["learned"] = learned_leg(ctx, jobs, frames, geom, B, args, fps)
        out["learned"]["speedup_vs_plain_graph"] = out["learned"]["value"] / gfps
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(frames, params, args.cpu_budget)
    ctx.close()
    if not args.no_mixed:
        out["mixed64"] = mixed_leg(args)
    if not args.no_large:
        out["large_motion"] = large_motion_leg(args)
    if args.copy_calibration_mb:
        _copy_calibration(args.copy_calibration_mb)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
