"""Benchmark: plain GPU motion estimation of 1080p B pictures with two
references (16x16 blocks at separation 12, 4-level pyramid, quarter pel,
lambda 10), batched, on one MI355X.

A step is one pass of the whole live ME sequence (pyramids and up-converted
references rebuilt, pel search, lambda map, sub-pel refinement, mode
decision, DC) over one batch of synthetic pictures resident in HBM.  Prints
one JSON line; the algorithmic byte and SAD-term counts behind the roofline
figures are in _geometry_terms().  The line also carries the learned-proposal
leg, the motion-compensation leg (residuals of the same batch from its
estimates, mc_leg()), the full-search leg (full_search_leg()) and the .imt
diagnostics-text leg (imt_leg()).
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
INT_LANE_OPS_PEAK = 256 * 64 * 2.4e9  # 256 CUs x 64 lanes per clock x 2.4 GHz


def _geometry_terms(geom, nrefs):
    """Algorithmic HBM bytes and SAD terms per picture for the main kernels.

    Bytes count what a kernel reads and writes per work unit with no reuse
    between units (int16 samples, int32 table entries); terms count the
    absolute differences of the SADs it evaluates.
    """
    xl, yl = geom.luma_blen[2][0], geom.luma_blen[2][1]
    B = geom.xnum_blocks * geom.ynum_blocks
    out = {}
    # pel tables: picture block + one or two reference regions of radius
    # R = r + 1 read, a 28-int record written, per block, level and reference
    b_bytes = t_terms = w_bytes = 0.0
    for l in range(geom.depth + 1):
        nb = geom.level_xblocks[l] * geom.level_yblocks[l]
        r = min(l + 1, 5)
        R = r + 1
        TS = (2 * R + 1) ** 2
        ntab = 2 if l < geom.depth else 1
        b_bytes += nb * nrefs * (2 * xl * yl + ntab * 2 * (xl + 2 * R) * (yl + 2 * R) + 8 + 28 * 4)
        t_terms += nb * nrefs * ntab * TS * xl * yl
        w_bytes += nb * nrefs * (28 * 4 + 8 + 12)
    out["k_pel_bulk"] = (b_bytes, t_terms)
    out["k_pel_wave"] = (w_bytes, 0.0)
    # sub-pel tables: picture block + 16 quarter-pel phase regions of
    # (xl + 2) x (yl + 2) bytes read, a 56-int record written
    out["k_subpel_bulk"] = (B * nrefs * (2 * xl * yl + 16 * (xl + 2) * (yl + 2) + 8 + 12 + 4 + 56 * 4),
                            B * nrefs * 49 * xl * yl)
    out["k_subpel_wave"] = (B * nrefs * (56 * 4 + 8 + 12), 0.0)
    # mode decision: level-1 and level-0 unit tables, superblock wavefront
    l1 = geom.luma_blen[1][0] * geom.luma_blen[1][1]
    l0 = geom.luma_blen[0][0] * geom.luma_blen[0][1]
    nu1, nsb = B // 4, B // 16
    out["k_mode_tables"] = (nu1 * (2 * l1 + nrefs * 4 * 4 * l1 + 24 * 4) + nsb * (2 * l0 + nrefs * 16 * 4 * l0 + 32 * 4),
                            nu1 * (nrefs * 4 + 16) * l1 + nsb * nrefs * 16 * l0)
    out["k_mode_wave"] = (nsb * (16 * 4 + 4 * 24 * 4 + 32 * 4 + 2 * l0 * 9 + 16 * 20), nsb * l0)
    out["k_upconvert"] = (nrefs * geom.width * geom.height * 2 * 5, 0.0)
    out["k_build_qp"] = (nrefs * (geom.width * geom.height * 4 * 2 + (geom.width + 2) * (geom.height + 2) * 16), 0.0)
    out["k_downconvert"] = (sum((1 + nrefs) * (geom.level_width[l] * geom.level_height[l] * 2 * 5)
                                for l in range(1, geom.depth + 1)), 0.0)
    return out


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


def cpu_baseline(frames, params, budget_s):
    """Reference encoder ME (oracle/_ref, single thread) or the oracle, on a
    bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O

    kind = "reference" if O.ref_available() else "port"
    run = O.run_ref if kind == "reference" else O.run_oracle
    n = 0
    t0 = time.time()
    while True:
        d = frames[n % len(frames)]
        run(params, d["cur"], d["refs"], d["u"], d["v"])
        n += 1
        if time.time() - t0 >= budget_s or n >= 200:
            break
    dt = time.time() - t0
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": n / dt, "unit": "frames/s", "cores": 1, "kind": kind,
            "sample": "%d 1080p B pictures (2 refs) of the same synthetic workload, %s %s, %.1f s on %s" % (
                n, "reference encoder ME (MMX build, compiled from its sources)" if kind == "reference" else "oracle port",
                "single thread", dt, cpu)}


def _mc_bytes(geom, me):
    """Algorithmic HBM bytes of compensating one picture (k_mc), counted per
    unit with no reuse between units, as the reference computes it
    (CompensateBlock per unit, mot_comp.cpp:541-603): each unit's OBMC
    window (clipped to the picture) reads 4 up-converted int16 samples per
    sample from each reference its mode uses (none for DC, two for
    bi-prediction); every sample of the three arrays is read and written once
    (int16); plus the MvData."""
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
        total += area_refs * 8 * (1 if c == 0 else 2) + pw * ph * 4 * (1 if c == 0 else 2)
    B = xnb * ynb
    return total + B * (2 * 8 + 4 + 3 * 2) + xsb * geom.ynum_sb * 4


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
           "int_ops_frac": 3.0 * terms / (ms * 1e-3) / INT_LANE_OPS_PEAK,
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
    return {"value": fps, "unit": "frames/s", "ms_per_step": dt / args.steps * 1e3,
            "speedup_vs_plain": fps / plain_fps, "mean_sb_cost_gap": float(np.mean(gaps)),
            "max_sb_cost_gap": float(np.max(gaps)), "fallback_rate": (f1 - f0) / max(1, b1 - b0),
            "pictures_for_gap": n, "weights": "dirac_me/proposal_weights.npy (2000 training pictures)"}


def mixed_leg(args):
    """The mixed batch of 64 pictures (16 CIF, 24 720p, 16 1080p, 8 2160p,
    1 or 2 references by seed, block parameters per size) in one
    dme_run_batch call per step: frames/s and luma samples/s."""
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
    ctx.close()
    return {"metric": "mixed-size ME frames/s (64 pictures per batch)", "value": len(jobs) / dt, "unit": "frames/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int16", "data": "synthetic",
            "config": {"workload": "mixed batch: 16 CIF, 24 1280x720, 16 1920x1080, 8 3840x2160, 1 or 2 refs by seed, "
                                   "blocks 12/8, 12/8, 16/12, 24/16, quarter pel, lambda 10", "batch": len(jobs)},
            "luma_samples_per_s": samples / dt,
            "kernel_ms_per_step": {k: round(v[0], 4) for k, v in kt.items()}}


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
    dom = max(kt, key=lambda k: kt[k][0])
    terms = _geometry_terms(geom, 2)
    alg_bytes, alg_terms = terms.get(dom, (0.0, 0.0))
    ms, nl = kt[dom]
    per_launch_ms = ms / nl
    achieved = alg_bytes * B / nl / (per_launch_ms * 1e-3) / 1e9
    roof = {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": _profiled_traffic(dom, B),
            "algorithmic_bytes_per_launch": alg_bytes * B / nl, "avg_launch_ms": per_launch_ms,
            "launches_per_step": nl}
    valu = {}
    for k in ("k_pel_bulk", "k_subpel_bulk", "k_mode_tables"):
        if k in kt:
            ops = 3.0 * terms[k][1] * B
            valu[k] = {"sad_terms_per_s": terms[k][1] * B / (kt[k][0] * 1e-3),
                       "int_ops_frac": ops / (kt[k][0] * 1e-3) / INT_LANE_OPS_PEAK}
    out = {
        "metric": METRIC, "value": fps, "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int16", "data": "synthetic",
        "config": {"workload": "plain GPU ME, 1920x1080 4:2:0 B pictures with 2 refs, blocks 16/12, 4-level "
                               "pyramid, quarter pel, lambda 10; batch of %d pictures per step" % B,
                   "batch": B, "width": 1920, "height": 1080, "refs": 2, "depth": 4, "precision": "quarter"},
        "block_decisions_per_s": fps * geom.xnum_blocks * geom.ynum_blocks * 2,
        "roofline": roof, "valu": valu,
        "kernel_ms_per_step": {k: round(v[0], 4) for k, v in kt.items()},
    }
    if not args.no_mc:
        out["motion_compensation"] = mc_leg(ctx, jobs, frames, params, geom, B, args)
    if not args.no_full:
        out["full_search"] = full_search_leg(ctx, jobs, frames, args)
    if not args.no_imt:
        out["imt"] = imt_leg(ctx, jobs, frames, geom, B, args)
    if not args.no_learned:
        out["learned"] = learned_leg(ctx, jobs, frames, geom, B, args, fps)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(frames, params, args.cpu_budget)
    ctx.close()
    if args.copy_calibration_mb:
        _copy_calibration(args.copy_calibration_mb)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
