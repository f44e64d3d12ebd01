"""GPU: the persistent form of the batch (PT_BATCH_PERSISTENT): one kernel
launch runs every job's device program.

* every job of a batch -- mixed sizes, patch radii, levels, exemplars,
  guides -- equals the same job synthesized alone, on repeated launches and
  for every claim size;
* the launch's grid holds only co-resident workgroups;
* a launch started with its give-up flag set stops and reports it, and the
  next launch is unaffected;
* jobs whose plans need a kernel without a device-program form are refused
  with the step named, before anything is launched.
"""
import pytest

from painttween_amd import Context, _ffi
from painttween_amd import workloads as W
from painttween_amd.forms import Batch
from painttween_amd.texsynth import PtError
from tests.test_gpu_forms import _batch_jobs, _same

pytestmark = pytest.mark.gpu


def _alone(jobs):
    with Context.gpu(0) as g:
        alone = [g.synthesize(g.default_params(**j["kw"]), j["frames"], j["style"], iterations=j["iterations"],
                              fixup_passes=0) for j in jobs]
        for j in jobs:
            j["params"] = g.default_params(**j["kw"])
    return alone


def _clear(b):
    for o in b.outputs:
        o.off.zero_()
        o.canvas.zero_()
        o.res.zero_()


def test_persistent_batch_equals_each_job_alone():
    jobs = _batch_jobs()
    alone = _alone(jobs)
    b = Batch(jobs)
    try:
        b.set(_ffi.PT_BATCH_PERSISTENT, 1)
        st = b.stats()
        assert st["persistent"] == 1 and st["persistent_tasks"] > 0 and st["persistent_tiles"] > 0, st
        # one workgroup per slot the device holds at once: at most 8 per CU on 256 CUs
        assert 0 < st["persistent_grid"] <= 8 * 256, st
        print("persistent grid %d, %d tasks, %d tiles, %d jobs on several XCDs" % (
            st["persistent_grid"], st["persistent_tasks"], st["persistent_tiles"], st["persistent_spanned"]), flush=True)
        # the 128x96 job holds about a third of the work: it is spread over several XCDs
        assert st["persistent_spanned"] >= 1, st
        for chunk in (8, 1, 16):
            b.set(_ffi.PT_BATCH_PERSIST_CHUNK, chunk)
            for rep in range(2):
                _clear(b)
                b.run()
                for i in range(len(jobs)):
                    _same(b.result(i), alone[i], "job %d, %d tiles per claim, run %d" % (i, chunk, rep))
    finally:
        b.close()


def test_persistent_single_job_spans_every_xcd():
    """One job alone holds the whole batch's work: its steps are split into
    one band per XCD, each band's writes written back once per step."""
    jobs = _batch_jobs()[2:3]
    alone = _alone(jobs)
    b = Batch(jobs)
    try:
        b.set(_ffi.PT_BATCH_PERSISTENT, 1)
        assert b.stats()["persistent_spanned"] == 1
        for rep in range(2):
            _clear(b)
            b.run()
            _same(b.result(0), alone[0], "single job, run %d" % rep)
    finally:
        b.close()


def test_persistent_give_up_flag_stops_the_launch_and_the_next_run_succeeds():
    jobs = _batch_jobs()[:2]
    alone = _alone(jobs)
    b = Batch(jobs)
    try:
        b.set(_ffi.PT_BATCH_PERSISTENT, 1)
        b.set(_ffi.PT_BATCH_PERSIST_ABORT, 1)
        with pytest.raises(PtError) as e:
            b.run()
        assert e.value.code == _ffi.PT_ERR_HIP and "gave up" in str(e.value)
        _clear(b)
        b.run()
        for i in range(len(jobs)):
            _same(b.result(i), alone[i], "job %d after a stopped launch" % i)
    finally:
        b.close()


def test_persistent_refuses_temporal_jobs_naming_the_step():
    st = W.style_pair(48, 40, seed=90, guides=True)
    frames = W.animation(40, 32, 2, seed=91)
    with Context.gpu(0) as g:
        p = g.default_params(residual_window_size=2, levels=2, temporal_coherence_weight=0.5)
    b = Batch([dict(params=p, frames=frames, style=st, iterations=2, scheme=_ffi.PT_S_FMBM_RM, fixup_passes=1)])
    try:
        b.set(_ffi.PT_BATCH_PERSISTENT, 1)
        with pytest.raises(PtError) as e:
            b.run()
        assert e.value.code == _ffi.PT_ERR_UNSUPPORTED and "persistent form" in str(e.value), str(e.value)
        # the graph form still runs the same batch
        b.set(_ffi.PT_BATCH_PERSISTENT, 0)
        b.run()
    finally:
        b.close()
