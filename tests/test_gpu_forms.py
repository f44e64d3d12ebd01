"""GPU: the streamed, batched and served forms.

* the per-frame (frame-major) path equals the oracle's run of the same
  schedule frame for frame, and its first frame is the reference's one-frame
  schedule (pt_synthesize, FMBM_RM, one pass), bit for bit;
* every job of a batch -- mixed sizes, patch radii, levels, exemplars --
  equals the same job synthesized alone, on eager runs, graph replays and
  after autotuning;
* every served frame equals the plain per-frame path, whether its sequence
  was served alone or mixed with others, and memory admission holds a
  sequence back until the budget frees up.
"""
import pytest

from painttween_amd import Context, _ffi
from painttween_amd import workloads as W
from painttween_amd.forms import Batch, DeviceImages, DeviceOutputs, Server, next_frame
from tests.oracle_ctx import OracleContext

pytestmark = pytest.mark.gpu


def _same(a, b, what):
    for name, x, y in zip(("offsets", "canvas", "residual"), a, b):
        assert x.tobytes() == y.tobytes(), "%s: %s differs" % (what, name)


def _stream(ctx, params, iterations, frames, style):
    """Plain per-frame path: one pt_synthesize_next_frame call per frame."""
    img = DeviceImages(0)
    sg = img.style(style)
    sh, sw = style["base"].shape[:2]
    h, w = frames[0]["color"].shape[:2]
    out = []
    for f in frames:
        o = DeviceOutputs(w, h, 1)
        next_frame(ctx, params, iterations, img.frame(f), sg, sw, sh, o)
        out.append(tuple(a[0] for a in o.numpy()))
    return out


def test_next_frame_matches_oracle():
    """The per-frame schedule, frame for frame, against the oracle's."""
    st = W.style_pair(64, 56, seed=41, guides=True)
    frames = W.animation(48, 40, 3, seed=42)
    kw = dict(residual_window_size=3, levels=3, temporal_coherence_weight=0.5)
    with Context.gpu(0) as g, OracleContext() as o:
        streamed = _stream(g, g.default_params(**kw), 2, frames, st)
        po = o.default_params(**kw)
        ref = [o.next_frame(po, 2, f, st) for f in frames]
    for f in range(len(frames)):
        _same(streamed[f], ref[f], "frame %d vs oracle" % f)


def test_next_frame_with_object_ids_matches_oracle():
    st = dict(W.style_pair(64, 56, seed=43, guides=True), object_ids=W.object_ids(43, 30, 64, 56))
    frames = [dict(f, object_id=W.object_ids(44, 31 + i, 48, 40)) for i, f in enumerate(W.animation(48, 40, 3, seed=44))]
    kw = dict(residual_window_size=3, levels=3, temporal_coherence_weight=0.5, style_obj_id_weight=0.6)
    with Context.gpu(0) as g, OracleContext() as o:
        streamed = _stream(g, g.default_params(**kw), 2, frames, st)
        po = o.default_params(**kw)
        ref = [o.next_frame(po, 2, f, st) for f in frames]
    for f in range(len(frames)):
        _same(streamed[f], ref[f], "frame %d vs oracle" % f)


def test_next_frame_with_luminance_shift_matches_oracle():
    st = W.style_pair(64, 56, seed=47, guides=True)
    frames = W.animation(48, 40, 3, seed=48)
    kw = dict(residual_window_size=3, levels=3, temporal_coherence_weight=0.5, lum_shift_enable=1)
    with Context.gpu(0) as g, OracleContext() as o:
        streamed = _stream(g, g.default_params(**kw), 2, frames, st)
        po = o.default_params(**kw)
        ref = [o.next_frame(po, 2, f, st) for f in frames]
    for f in range(len(frames)):
        _same(streamed[f], ref[f], "frame %d vs oracle" % f)
    assert (streamed[2][0]["lum_y"] != 0).any()


def test_first_streamed_frame_is_the_still_schedule():
    """Frame 0 of a sequence runs exactly the one-frame reference schedule."""
    st = W.style_pair(64, 56, seed=45, guides=True)
    frames = W.animation(48, 40, 1, seed=46)
    with Context.gpu(0) as g:
        p = g.default_params(residual_window_size=3, levels=3)
        streamed = _stream(g, p, 2, frames, st)
        lm = g.synthesize(p, frames, st, scheme=_ffi.PT_S_FMBM_RM, iterations=2, fixup_passes=0)
    _same(streamed[0], tuple(a[0] for a in lm), "frame 0 vs one-frame schedule")


def test_sequence_reset_restarts():
    st = W.style_pair(48, 48, seed=43)
    frames = W.animation(40, 32, 2, seed=44)
    with Context.gpu(0) as g:
        p = g.default_params(residual_window_size=2, levels=2)
        a = _stream(g, p, 2, frames, st)
        g._check(g.lib.pt_sequence_reset(g.handle))
        b = _stream(g, p, 2, frames, st)
    for x, y in zip(a, b):
        _same(x, y, "after reset")


def _batch_jobs():
    """Uneven jobs: sizes, aspect, patch radius, levels, exemplars, guides."""
    specs = [(96, 64, 2, 3, 64, 64, True), (40, 56, 3, 2, 48, 40, False), (128, 96, 3, 4, 96, 80, True),
             (72, 72, 2, 3, 56, 64, True), (33, 47, 3, 2, 40, 36, True)]
    jobs = []
    for i, (w, h, r, lv, sw, sh, guides) in enumerate(specs):
        st = W.style_pair(sw, sh, seed=50 + i, guides=guides)
        fr = W.still_frame(w, h, seed=60 + i, guides=guides)
        jobs.append(dict(kw=dict(residual_window_size=r, levels=lv, dist_trans_weight=0.5 if guides else 0.0,
                                 orientation_enabled=int(guides)),
                         frames=[fr], style=st, iterations=2))
    return jobs


def test_batch_equals_each_job_alone():
    jobs = _batch_jobs()
    with Context.gpu(0) as g:
        alone = []
        for j in jobs:
            alone.append(g.synthesize(g.default_params(**j["kw"]), j["frames"], j["style"], iterations=2,
                                      fixup_passes=0))
        for j in jobs:
            j["params"] = g.default_params(**j["kw"])
    b = Batch(jobs)
    try:
        st = b.stats()
        assert st["multi_launches"] < st["ops"] / 2, st   # steps of different jobs share launches
        b.run()                                           # captures the graph
        for i in range(len(jobs)):
            _same(b.result(i), alone[i], "job %d graph run 1" % i)
        for i in range(len(jobs)):                        # clear outputs, replay
            b.outputs[i].off.zero_()
            b.outputs[i].canvas.zero_()
            b.outputs[i].res.zero_()
        b.run()
        for i in range(len(jobs)):
            _same(b.result(i), alone[i], "job %d graph replay" % i)
        b.set(_ffi.PT_BATCH_USE_GRAPH, 0)
        b.set(_ffi.PT_BATCH_XCD_REMAP, 0)
        b.run()
        for i in range(len(jobs)):
            _same(b.result(i), alone[i], "job %d eager" % i)
        b.set(_ffi.PT_BATCH_USE_GRAPH, 1)
        b.set(_ffi.PT_BATCH_XCD_REMAP, 1)
        b.tune(1)
        b.run()
        for i in range(len(jobs)):
            _same(b.result(i), alone[i], "job %d tuned" % i)
    finally:
        b.close()


@pytest.mark.parametrize("shape", range(7))
def test_batch_every_shape_equals_alone(shape):
    """Each kernel shape, forced on every level class, gives the same jobs."""
    jobs = _batch_jobs()[:3]
    with Context.gpu(0) as g:
        alone = [g.synthesize(g.default_params(**j["kw"]), j["frames"], j["style"], iterations=2, fixup_passes=0)
                 for j in jobs]
        for j in jobs:
            j["params"] = g.default_params(**j["kw"])
    b = Batch(jobs)
    try:
        for op in (0, 1):
            for cls in range(3):
                b.set(_ffi.PT_BATCH_SHAPE, op * 256 + cls * 16 + shape)
        b.run()
        for i in range(len(jobs)):
            _same(b.result(i), alone[i], "job %d shape %d" % (i, shape))
    finally:
        b.close()


def test_batch_of_every_schedule_equals_each_job_alone():
    """Jobs of different schedules (FMBM, FMBM_RM, IND, FBMO, TCTF) planned
    together, graph-replayed: each equals the same job synthesized alone."""
    st = W.style_pair(48, 40, seed=90, guides=True)
    frames = W.animation(40, 32, 3, seed=91)
    jobs = []
    for scheme in (_ffi.PT_S_FMBM, _ffi.PT_S_FMBM_RM, _ffi.PT_S_IND, _ffi.PT_S_FBMO, _ffi.PT_S_TCTF):
        jobs.append(dict(kw=dict(residual_window_size=2, levels=2, temporal_coherence_weight=0.5),
                         frames=frames, style=st, iterations=2, scheme=scheme, fixup_passes=1))
    with Context.gpu(0) as g:
        alone = [g.synthesize(g.default_params(**j["kw"]), j["frames"], j["style"], scheme=j["scheme"], iterations=2,
                              fixup_passes=1) for j in jobs]
        for j in jobs:
            j["params"] = g.default_params(**j["kw"])
    b = Batch(jobs)
    try:
        b.run()
        for i in range(len(jobs)):
            _same(b.result(i), alone[i], "job %d (scheme %d)" % (i, jobs[i]["scheme"]))
    finally:
        b.close()


def test_batch_mixing_object_id_jobs_equals_each_alone():
    """A job with the object-ID term (extended-terms kernels) and one without
    planned together: each equals the same job synthesized alone."""
    st = W.style_pair(48, 40, seed=92, guides=True)
    frames = W.animation(40, 32, 2, seed=93)
    st_ids = dict(st, object_ids=W.object_ids(92, 30, 48, 40))
    fr_ids = [dict(f, object_id=W.object_ids(93, 31 + i, 40, 32)) for i, f in enumerate(frames)]
    jobs = [dict(kw=dict(residual_window_size=2, levels=2, style_obj_id_weight=0.6), frames=fr_ids, style=st_ids),
            dict(kw=dict(residual_window_size=2, levels=2), frames=frames, style=st)]
    for j in jobs:
        j.update(iterations=2, scheme=_ffi.PT_S_FMBM_RM, fixup_passes=1)
    with Context.gpu(0) as g:
        alone = [g.synthesize(g.default_params(**j["kw"]), j["frames"], j["style"], iterations=2) for j in jobs]
        for j in jobs:
            j["params"] = g.default_params(**j["kw"])
    b = Batch(jobs)
    try:
        b.run()
        for i in range(len(jobs)):
            _same(b.result(i), alone[i], "job %d" % i)
    finally:
        b.close()
    assert alone[0][0].tobytes() != alone[1][0].tobytes()



def test_batch_with_max_level_and_backward_first_pass_equals_each_alone():
    """maxLevel and the first pass direction travel in each job's parameters:
    a maxLevel-1 FMBM job, a backward-first FMBM_RM job and a plain job
    planned together each equal the same job synthesized alone."""
    st = W.style_pair(48, 40, seed=94, guides=True)
    frames = W.animation(40, 32, 3, seed=95)
    base = dict(residual_window_size=2, levels=3, temporal_coherence_weight=0.5)
    jobs = [dict(kw=dict(base, max_level=1), scheme=_ffi.PT_S_FMBM),
            dict(kw=dict(base, first_pass_direction=1), scheme=_ffi.PT_S_FMBM_RM),
            dict(kw=dict(base), scheme=_ffi.PT_S_FMBM_RM)]
    for j in jobs:
        j.update(iterations=2, fixup_passes=1, frames=frames, style=st)
    with Context.gpu(0) as g:
        alone = [g.synthesize(g.default_params(**j["kw"]), frames, st, scheme=j["scheme"], iterations=2,
                              fixup_passes=1) for j in jobs]
        for j in jobs:
            j["params"] = g.default_params(**j["kw"])
    b = Batch(jobs)
    try:
        b.run()
        for i in range(len(jobs)):
            _same(b.result(i), alone[i], "job %d" % i)
    finally:
        b.close()
    assert alone[0][0].tobytes() != alone[2][0].tobytes() and alone[1][0].tobytes() != alone[2][0].tobytes()

def _sequences():
    seqs = []
    for i, (w, h, sw, sh) in enumerate([(48, 40, 64, 56), (64, 48, 48, 48), (40, 40, 56, 64)]):
        seqs.append(dict(style=W.style_pair(sw, sh, seed=70 + i, guides=True),
                         frames=W.animation(w, h, 3, seed=80 + i), w=w, h=h,
                         kw=dict(residual_window_size=3 if i != 1 else 2, levels=3, temporal_coherence_weight=0.5)))
    return seqs


def test_served_frames_equal_plain_alone_and_mixed():
    seqs = _sequences()
    with Context.gpu(0) as g:
        plain = []
        for s in seqs:
            with Context.gpu(0) as c:
                plain.append(_stream(c, c.default_params(**s["kw"]), 2, s["frames"], s["style"]))
        params = [g.default_params(**s["kw"]) for s in seqs]
    # each sequence served alone
    for k, s in enumerate(seqs):
        srv = Server(8 << 30, 4)
        try:
            q = srv.open(params[k], 2, s["w"], s["h"], s["style"])
            tickets = [srv.submit(q, f) for f in s["frames"]]
            srv.drain()
            for f, t in enumerate(tickets):
                _same(tuple(a[0] for a in srv.result(t)), plain[k][f], "seq %d frame %d alone" % (k, f))
        finally:
            srv.close()
    # all sequences mixed: frames interleaved, several in flight
    srv = Server(8 << 30, 3)
    try:
        qs = [srv.open(params[k], 2, s["w"], s["h"], s["style"]) for k, s in enumerate(seqs)]
        tickets = {}
        for f in range(3):
            for k, s in enumerate(seqs):
                tickets[(k, f)] = srv.submit(qs[k], s["frames"][f])
            srv.poll(max_ticks=25)       # later frames arrive while earlier ones run
        srv.drain()
        st = srv.stats()
        assert st["completed"] == 9 and st["multi_launches"] < st["launches"]
        for (k, f), t in tickets.items():
            _same(tuple(a[0] for a in srv.result(t)), plain[k][f], "seq %d frame %d mixed" % (k, f))
    finally:
        srv.close()


def test_admission_waits_for_budget():
    seqs = _sequences()[:2]
    with Context.gpu(0) as g:
        params = [g.default_params(**s["kw"]) for s in seqs]
    # a budget that holds either sequence but not both
    budget = int(1.5 * max(_seq_bytes(s) for s in seqs))
    assert budget < _seq_bytes(seqs[0]) + _seq_bytes(seqs[1])
    srv = Server(budget, 2)
    try:
        a = srv.open(params[0], 2, seqs[0]["w"], seqs[0]["h"], seqs[0]["style"])
        b = srv.open(params[1], 2, seqs[1]["w"], seqs[1]["h"], seqs[1]["style"])
        ta = srv.submit(a, seqs[0]["frames"][0])
        tb = srv.submit(b, seqs[1]["frames"][0])
        srv.poll(max_ticks=5)
        st = srv.stats()
        assert st["resident_sequences"] == 1 and st["waiting_frames"] == 1, st
        done = set()
        for _ in range(1000):
            done |= {t for t, _ in srv.poll(max_ticks=50)}
            if ta in done:
                break
        assert ta in done and tb not in done
        srv.close_sequence(a)          # frees the budget: b is admitted
        srv.drain()
        assert srv.stats()["completed"] == 2
    finally:
        srv.close()


def _seq_bytes(s):
    sh, sw = s["style"]["base"].shape[:2]
    # pt_sequence_bytes estimate, mirrored: 344 B/px working buffers, 32 B per tap and pixel of the forward
    # time table, 2*60*4/3 + 2*48*4/3 B/px pyramids and snapshots, 40 + 52*4/3 B/texel, +10%, +8 MiB
    px, ex = s["w"] * s["h"], sw * sh
    taps = (2 * s["kw"]["residual_window_size"] + 1) ** 2
    tt = 32 * taps if taps <= 81 else 0
    return int((px * (344 + tt + 2 * 60 * 4 / 3 + 2 * 48 * 4 / 3) + ex * (40 + 52 * 4 / 3)) * 1.1) + (8 << 20)
