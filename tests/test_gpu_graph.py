"""Launch tuning and the captured batch step (dme_set_tuning,
dme_autotune_batch, dme_tuning_save/load, dme_batch_graph_*).

No launch setting may change a result: a mixed batch (pictures of three
sizes, one and two references, pictures shared between jobs, half, quarter
and eighth pel) gives the same MEData bytes under every setting.  A graph
replay gives the bytes of dme_run_batch, replay after replay with eager
calls of the same context between them, reads the current contents of
device-bound pictures, and refuses to run once a picture it reads was
registered again.  The eager results themselves are pinned to the oracle by
the batch parity tests."""
import numpy as np
import pytest

import dirac_me
from dirac_me import abi, make_params, synth
from dirac_me.learned import DeviceArray

pytestmark = pytest.mark.gpu


def _assert_same(want, got, nrefs, tag=""):
    for k, v in want.arrays(nrefs).items():
        g = got.arrays(nrefs)[k]
        if v.tobytes() != g.tobytes():
            n = int((v.view(np.uint8) != g.view(np.uint8)).sum())
            raise AssertionError("%s %s differs (%d bytes)" % (tag, k, n))


@pytest.fixture(scope="module")
def gctx():
    c = dirac_me.Context(0)
    yield c
    c.close()


def _mixed(ctx, base):
    """Seven jobs: CIF (2 refs, quarter; 1 ref, eighth), 720p (2 refs, half),
    1080p (2 refs, quarter) and three jobs sharing one picture set (a
    picture is the reference of another job)."""
    specs = [("cif", 2, 11, 2), ("cif", 1, 12, 3), ("720p", 2, 13, 1), ("1080p", 2, 14, 2)]
    jobs, geoms = [], []
    for j, (size, nrefs, seed, prec) in enumerate(specs):
        d = synth.sized(seed, size, nrefs)
        p = make_params(*d["block"], precision=prec, depth=d["depth"], lam=10.0)
        b = base + 3 * j
        nums = [b, b + 2][:nrefs]
        for n, r in zip(nums, d["refs"]):
            ctx.upload(n, r)
        ctx.upload(b + 1, d["cur"], d["u"], d["v"])
        jobs.append((b + 1, nums, p))
        H, W = d["cur"].shape
        geoms.append(dirac_me.geometry(p, W, H))
    # shared: pictures s0..s3 of one CIF sequence; s1 refs (s0, s2), s2 refs
    # (s1, s3), s3 refs s2
    d = synth.sized(15, "cif", 2)
    seq = [d["refs"][0], d["cur"], d["refs"][1], synth.sized(16, "cif", 1)["cur"]]
    s = base + 3 * len(specs)
    for k, pic in enumerate(seq):
        ctx.upload(s + k, pic, *synth.chroma(pic, 40 + k))
    p = make_params(*d["block"], precision=2, depth=d["depth"], lam=10.0)
    for cur, refs in ((s + 1, [s, s + 2]), (s + 2, [s + 1, s + 3]), (s + 3, [s + 2])):
        jobs.append((cur, refs, p))
        geoms.append(dirac_me.geometry(p, 352, 288))
    return jobs, geoms


def _run(ctx, jobs, geoms):
    outs = [dirac_me.MEData(g) for g in geoms]
    ctx.invalidate(-1)
    ctx.run_batch(jobs, outs)
    return outs


SETTINGS = [
    dict(parts=1, stagger=0),
    dict(parts=2, stagger=0),
    dict(parts=3, stagger=1),
    dict(parts=4, stagger=0),
    dict(parts=4, stagger=1, pel_wave_threads=256, subpel_wave_threads=256, mode_wave_threads=256),
    dict(parts=2, stagger=1, pel_wave_threads=512, subpel_wave_threads=512, mode_wave_threads=512),
]


def test_launch_settings_do_not_change_results(gctx):
    jobs, geoms = _mixed(gctx, 8000)
    gctx.reset_tuning()
    want = _run(gctx, jobs, geoms)
    for st in SETTINGS:
        print("setting", st, flush=True)
        gctx.set_tuning(**st)
        got = _run(gctx, jobs, geoms)
        for (pn, refs, _), w, g in zip(jobs, want, got):
            _assert_same(w, g, len(refs), "%s picture %d" % (st, pn))
    gctx.reset_tuning()
    t = gctx.get_tuning()
    assert (t.parts, t.stagger, t.pel_wave_threads, t.subpel_wave_threads, t.mode_wave_threads) == (4, 1, 1024, 512,
                                                                                                    1024)
    with pytest.raises(RuntimeError):
        gctx.set_tuning(parts=5)
    with pytest.raises(RuntimeError):
        gctx.set_tuning(subpel_wave_threads=1024)


def test_graph_replays_equal_the_eager_batch(gctx):
    jobs, geoms = _mixed(gctx, 8100)
    want = _run(gctx, jobs, geoms)
    other, ogeoms = jobs[:2], geoms[:2]
    g = gctx.batch_graph(jobs)
    try:
        for rep in range(4):
            print("replay", rep, flush=True)
            got = [dirac_me.MEData(x) for x in geoms]
            g.launch(got)
            for (pn, refs, _), w, o in zip(jobs, want, got):
                _assert_same(w, o, len(refs), "replay %d picture %d" % (rep, pn))
            # eager work of the same context between replays
            _run(gctx, other, ogeoms)
            gctx.synchronize()
        g.launch()  # outputs may stay on the device
    finally:
        g.close()


def test_graph_reads_the_current_contents_of_bound_pictures(gctx):
    a, b = synth.sized(21, "720p", 2), synth.sized(22, "720p", 2)
    H, W = a["cur"].shape
    p = make_params(*a["block"], precision=2, depth=a["depth"], lam=10.0)
    geom = dirac_me.geometry(p, W, H)
    planes = {}
    for k, (y, u, v) in enumerate([(a["refs"][0], None, None), (a["cur"], a["u"], a["v"]), (a["refs"][1], None, None)]):
        dy = DeviceArray(gctx, (H, W), np.int16)
        dy.upload(y)
        du = dv = None
        if u is not None:
            du, dv = DeviceArray(gctx, (H // 2, W // 2), np.int16), DeviceArray(gctx, (H // 2, W // 2), np.int16)
            du.upload(u)
            dv.upload(v)
        planes[8200 + k] = (dy, du, dv)
        gctx.bind_device(8200 + k, W, H, dy.at(), du.at() if du else None, dv.at() if dv else None)
    jobs = [(8201, [8200, 8202], p)]
    g = gctx.batch_graph(jobs)
    try:
        for d in (a, b, a):
            planes[8200][0].upload(d["refs"][0])
            planes[8201][0].upload(d["cur"])
            planes[8201][1].upload(d["u"])
            planes[8201][2].upload(d["v"])
            planes[8202][0].upload(d["refs"][1])
            want = _run(gctx, jobs, [geom])[0]
            got = dirac_me.MEData(geom)
            g.launch([got])
            _assert_same(want, got, 2, "bound pictures")
        # a picture registered again: the graph refuses to run
        gctx.upload(8200, b["refs"][0])
        with pytest.raises(RuntimeError, match="registered again"):
            g.launch()
    finally:
        g.close()
        for n in (8200, 8201, 8202):
            gctx.drop(n)


def test_autotune_table_and_its_file(gctx, tmp_path):
    jobs, geoms = _mixed(gctx, 8300)
    want = _run(gctx, jobs, geoms)
    t = gctx.autotune_batch(jobs, reps=1)
    print("tuned", t.parts, t.stagger, t.pel_wave_threads, t.subpel_wave_threads, t.mode_wave_threads,
          "%.2f ms" % t.step_ms, flush=True)
    assert 1 <= t.parts <= 4 and t.step_ms > 0
    got = _run(gctx, jobs, geoms)  # runs with the table's setting
    for (pn, refs, _), w, o in zip(jobs, want, got):
        _assert_same(w, o, len(refs), "tuned picture %d" % pn)
    path = str(tmp_path / "tuning.txt")
    gctx.save_tuning(path)
    text = open(path).read()
    assert text.startswith("#") and len(text.strip().splitlines()) == 2
    c2 = dirac_me.Context(0)
    try:
        c2.load_tuning(path)
        c2.save_tuning(path + ".2")
        assert open(path + ".2").read() == text
        bad = tmp_path / "bad.txt"
        bad.write_text("x 9 1 1024 512 1024 1.0\n")
        with pytest.raises(RuntimeError, match="bad tuning line 1"):
            c2.load_tuning(str(bad))
    finally:
        c2.close()
    # a graph takes the table's setting too
    g = gctx.batch_graph(jobs)
    try:
        got = [dirac_me.MEData(x) for x in geoms]
        g.launch(got)
        for (pn, refs, _), w, o in zip(jobs, want, got):
            _assert_same(w, o, len(refs), "tuned graph picture %d" % pn)
    finally:
        g.close()
