"""GPU tests at the benchmark's own workloads (not reduced sizes):

* the 1024^2 still (exemplar pair 1024^2, 7x7 patches, 5 levels, 8
  iterations, orientation and distance-transform guides): the whole schedule,
  level 0 included, bit for bit against the oracle;
* the 960x540 animation (exemplar 512^2, 7x7 patches, 4 levels, 3
  iterations, two passes, temporal coherence 0.5): three frames bit for bit
  against the oracle, and all 24 frames under size-independent checks (the
  canvas recomputed by the oracle from the GPU offsets of every frame, a
  second run reproducing the first, offsets inside the exemplar);
* the batched workload (64 jobs, 128^2..1024^2, aspect 0.5..2, 5x5/7x7, 3-5
  levels, 6 iterations): every job of the tuned batch equals the job run
  alone on the plain path;
* the served workload (32 sequences, 256..1920 x 256..1080, exemplars
  512..1024, 3 frames each): every served frame equals the plain per-frame
  path, served alone and mixed.
"""
from multiprocessing.pool import ThreadPool

import numpy as np
import pytest

from painttween_amd import Context, _ffi
from painttween_amd import workloads as W
from painttween_amd.forms import Batch, Server
from tests.harness import Session, first_difference
from tests.oracle_ctx import OracleContext
from tests.test_gpu_forms import _same, _stream

pytestmark = pytest.mark.gpu

STILL_KW = dict(residual_window_size=3, levels=5, dist_trans_weight=0.5, input_analogy_weight=1.0,
                orientation_enabled=1, canvas_output_weight=1.0)
ANIM_KW = dict(residual_window_size=3, levels=4, temporal_coherence_weight=0.5, dist_trans_weight=0.5,
               input_analogy_weight=1.0, orientation_enabled=1, canvas_output_weight=1.0)


def _cmp(g_out, o_out, what):
    for label, a, b in zip(("offsets", "canvas", "residual"), g_out, o_out):
        d = first_difference(a, b)
        assert d is None, "%s: %s %s" % (what, label, d)


def test_headline_still_whole_schedule_bit_exact():
    """Level 0 holds three quarters of the patch distances and is the only
    level where the default policy takes the large-level kernels (exemplar
    window, isolated big launches): the final offsets, canvas and residual of
    the full schedule equal the oracle's."""
    st = W.style_pair(1024, 1024, seed=2, guides=True)
    fr = W.still_frame(1024, 1024, seed=2, guides=True)
    with Context.gpu(0) as g:
        a = g.synthesize(g.default_params(**STILL_KW), [fr], st, iterations=8)
    with OracleContext() as o:
        b = o.synthesize(o.default_params(**STILL_KW), [fr], st, iterations=8)
    _cmp(a, b, "1024^2 still")


def _anim_inputs(frames):
    return W.animation(960, 540, frames, seed=3), W.style_pair(512, 512, seed=3, guides=True)


def test_animation_workload_three_frames_bit_exact():
    """960x540 with the odd level chain 480x270, 240x135, 120x67; both passes
    (forward then backward advection) of three frames against the oracle."""
    frames, st = _anim_inputs(3)
    with Context.gpu(0) as g:
        a = g.synthesize(g.default_params(**ANIM_KW), frames, st, iterations=3, fixup_passes=1)
    with OracleContext() as o:
        b = o.synthesize(o.default_params(**ANIM_KW), frames, st, iterations=3, fixup_passes=1)
    _cmp(a, b, "960x540 animation, 3 frames")


def test_animation_workload_24_frames_properties():
    """All 24 frames: a second run reproduces the first bit for bit, offsets
    lie inside the exemplar, energies are finite and non-negative, and the
    oracle, given each frame's GPU offsets, rebuilds its canvas bit for bit
    (offsetsToOutput: one bilinear exemplar-output fetch per pixel)."""
    frames, st = _anim_inputs(24)
    with Context.gpu(0) as g:
        p = g.default_params(**ANIM_KW)
        a = g.synthesize(p, frames, st, iterations=3, fixup_passes=1)
        b = g.synthesize(p, frames, st, iterations=3, fixup_passes=1)
    _cmp(a, b, "run to run")
    off, canvas, res = a
    assert off.shape == (24, 540, 960)
    for f in ("x", "y"):
        assert off[f].min() >= 0.0 and off[f].max() <= 513.0
    assert np.isfinite(res).all() and (res >= 0).all()
    with OracleContext() as o:
        s = Session(o, st, [frames[0]], o.default_params(**ANIM_KW))
        s.run([("style", 0), ("frame", 0, 0)])
        for k in range(24):
            o.load_offsets(0, off[k])
            o.sync_buffers(0, _ffi.PT_DONT_COPY_OFFSETS | _ffi.PT_DONT_UPDATE_RESIDUAL)
            _, oc, _ = s.state(0)
            assert first_difference(canvas[k], oc) is None, "frame %d canvas" % k


def test_batched_workload_64_jobs_equal_plain_after_tuning():
    specs = W.batch_specs(64)
    with ThreadPool(8) as pool:
        inputs = pool.map(W.batch_job, specs)
    jobs = []
    with Context.gpu(0) as g:
        alone = []
        for spec, (fr, st) in zip(specs, inputs):
            kw = dict(residual_window_size=spec["r"], levels=spec["levels"], dist_trans_weight=0.5,
                      input_analogy_weight=1.0, orientation_enabled=1, canvas_output_weight=1.0)
            p = g.default_params(**kw)
            alone.append(g.synthesize(p, [fr], st, iterations=spec["iterations"], fixup_passes=0))
            jobs.append(dict(params=p, frames=[fr], style=st, iterations=spec["iterations"]))
    b = Batch(jobs)
    try:
        b.tune(1)
        b.run()
        st = b.stats()
        assert st["jobs"] == 64 and st["multi_launches"] > 0
        for i in range(64):
            _same(b.result(i), alone[i], "job %d (%dx%d, r=%d, %d levels)" % (
                i, specs[i]["w"], specs[i]["h"], specs[i]["r"], specs[i]["levels"]))
        # the persistent form: the whole batch in one launch, claim size tuned
        for o in b.outputs:
            o.off.zero_()
            o.canvas.zero_()
            o.res.zero_()
        b.set(_ffi.PT_BATCH_PERSISTENT, 1)
        b.tune(1)
        b.run()
        st = b.stats()
        assert st["persistent"] == 1 and st["persistent_tasks"] > 0
        for i in range(64):
            _same(b.result(i), alone[i], "persistent job %d (%dx%d, r=%d, %d levels)" % (
                i, specs[i]["w"], specs[i]["h"], specs[i]["r"], specs[i]["levels"]))
    finally:
        b.close()


def test_served_workload_32_sequences_equal_plain_alone_and_mixed():
    specs = W.serving_specs(32)
    with ThreadPool(8) as pool:
        seqs = pool.map(lambda s: W.serving_sequence(s, 3), specs)
    kw = dict(ANIM_KW)
    plain = []
    for (frames, st) in seqs:
        with Context.gpu(0) as c:
            plain.append(_stream(c, c.default_params(**kw), 3, frames, st))
    with Context.gpu(0) as g:
        params = g.default_params(**kw)
    # four sequences served each on its own server
    for k in (0, 7, 19, 31):
        frames, st = seqs[k]
        srv = Server(200 << 30, 4)
        try:
            q = srv.open(params, 3, specs[k]["w"], specs[k]["h"], st)
            tickets = [srv.submit(q, f) for f in frames]
            srv.drain()
            for f, t in enumerate(tickets):
                _same(tuple(a[0] for a in srv.result(t)), plain[k][f], "seq %d frame %d alone" % (k, f))
        finally:
            srv.close()
    # all 32 mixed under the 200 GB budget: frames of a round arrive while
    # the previous round runs
    srv = Server(200 << 30, 32)
    try:
        qs = [srv.open(params, 3, specs[k]["w"], specs[k]["h"], seqs[k][1]) for k in range(32)]
        tickets = {}
        for f in range(3):
            for k in range(32):
                tickets[(k, f)] = srv.submit(qs[k], seqs[k][0][f])
            srv.poll(max_ticks=40)
        srv.drain()
        stt = srv.stats()
        assert stt["completed"] == 96 and stt["multi_launches"] > 0
        for (k, f), t in tickets.items():
            _same(tuple(a[0] for a in srv.result(t)), plain[k][f], "seq %d frame %d mixed" % (k, f))
    finally:
        srv.close()
