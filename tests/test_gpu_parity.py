"""GPU parity: the HIP library against the CPU oracle, bit for bit, through
the C ABI.  Offsets, canvas and residual are compared as raw bytes after every
operator (the offset field is integer-exact work by construction; canvas and
residual are FP32 and required identical too, which is stricter than the
1e-5 relative tolerance the target allows)."""
import numpy as np
import pytest

from painttween_amd import Context, _ffi
from painttween_amd import workloads as W
from tests.harness import Session, assert_same_states, first_difference
from tests.oracle_ctx import OracleContext

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    ctx = Context.gpu(0)
    yield ctx
    ctx.close()


def _both(style, frames, overrides, ops, early_exit=1, unit_parallel=1, tap=0, time_table=None):
    """Run `ops` on the GPU and on the oracle.  tap=N > 0 sends every
    propagate and residual refresh with a supported radius to the
    tap-parallel kernels with N tap evaluations per workgroup; tap=0 keeps
    them off so the pixel and unit-parallel kernels run.  time_table=0
    turns the per-level tables of the temporal terms off."""
    g = Context.gpu(0)
    if time_table is not None:
        g.set_tuning(_ffi.PT_TUNE_TIME_TABLE_MAX_TAPS, time_table)
    g.set_tuning(_ffi.PT_TUNE_EARLY_EXIT, early_exit)
    g.set_tuning(_ffi.PT_TUNE_UNIT_PARALLEL, unit_parallel)
    g.set_tuning(_ffi.PT_TUNE_EARLY_EXIT_MIN_PIXELS, 0)  # exercise the early-exit kernels on small levels
    g.set_tuning(_ffi.PT_TUNE_TAP_PARALLEL_MAX_PIXELS, (1 << 30) if tap else 0)
    if tap:
        g.set_tuning(_ffi.PT_TUNE_TAP_PARALLEL_EVALS, tap)
    o = OracleContext()
    try:
        sg = Session(g, style, frames, g.default_params(**overrides))
        so = Session(o, style, frames, o.default_params(**overrides))
        return sg.run(ops), so.run(ops)
    finally:
        g.close()
        o.close()


def test_pyramids_bit_exact():
    st = W.style_pair(48, 40, seed=3, guides=True)
    fr = W.animation(40, 36, 2, seed=4)[1]
    with Context.gpu(0) as g, OracleContext() as o:
        cases = [(_ffi.PT_ELEM_F32X4, st["base"], 1), (_ffi.PT_ELEM_F32X3, st["orientation"], 1),
                 (_ffi.PT_ELEM_F32, st["dist_trans"], 1), (_ffi.PT_ELEM_F32X2, fr["vel_b"], 0)]
        for elem, arr, mode in cases:
            pg, po = g.pyramid(elem, arr, mode), o.pyramid(elem, arr, mode)
            assert pg.levels == po.levels
            for l in range(pg.levels):
                assert first_difference(pg.download(l), po.download(l)) is None, (elem, l)
        tg, to = g.pyramid(3, fr["orientation"]), o.pyramid(3, fr["orientation"])
        ag, ao = g.orientation_pyramid(tg), o.orientation_pyramid(to)
        for l in range(ag.levels):
            assert first_difference(ag.download(l), ao.download(l)) is None, l


def _still_ops(levels, iterations=2):
    ops = []
    for lv in range(levels - 1, -1, -1):
        ops += [("style", lv), ("reset_adv",), ("frame", lv, 0)]
        if lv == levels - 1:
            ops += [("randomize",)]
        else:
            ops += [("upsample",), ("sync", 0)]
        ops += [("sync", 0)]
        for it in range(iterations):
            ops += [("propagate", j) for j in (128, 64, 32, 16, 8, 4, 2, 1, 1)]
            if it < iterations - 1:
                ops += [("scatter",)]
    return ops


@pytest.mark.parametrize("early_exit,unit_parallel,tap", [(1, 1, 0), (0, 1, 0), (1, 0, 0), (0, 0, 0), (1, 1, 1024),
                                                         (1, 1, 300)])
@pytest.mark.parametrize("r,orient,dtw", [(2, False, 0.0), (3, True, 0.5), (1, True, 0.0), (4, False, 0.0),
                                          (5, True, 0.0)])
def test_still_operator_sequence(r, orient, dtw, early_exit, unit_parallel, tap):
    st = W.style_pair(40, 36, seed=5, guides=orient or dtw > 0)
    fr = W.still_frame(44, 34, seed=6, guides=orient or dtw > 0)
    ov = dict(residual_window_size=r, levels=3, orientation_enabled=int(orient), dist_trans_weight=dtw)
    ga, oa = _both(st, [fr], ov, _still_ops(3), early_exit, unit_parallel, tap)
    assert_same_states(ga, oa)


@pytest.mark.parametrize("early_exit,unit_parallel,tap", [(1, 1, 0), (0, 1, 0), (1, 0, 0), (1, 1, 1024)])
def test_counter_rng_scatter(early_exit, unit_parallel, tap):
    st = W.style_pair(64, 64, seed=7)
    fr = W.still_frame(48, 40, seed=8)
    ov = dict(residual_window_size=2, levels=2, rng_mode=_ffi.PT_RNG_COUNTER, seed=1234, orientation_enabled=0)
    ga, oa = _both(st, [fr], ov, _still_ops(2, iterations=3), early_exit, unit_parallel, tap)
    assert_same_states(ga, oa)


@pytest.mark.parametrize("tap", [0, 1024, 1])
def test_odd_sizes_and_transparency(tap):
    st = W.style_pair(37, 29, seed=9)
    st["output"][..., 3] = np.where(st["output"][..., 3] > 0.6, 1.0, 0.2).astype(np.float32)
    fr = W.still_frame(45, 27, seed=10)
    fr["color"][..., 3] = 1.0
    ov = dict(residual_window_size=2, levels=3)
    ga, oa = _both(st, [fr], ov, _still_ops(3), tap=tap)
    assert_same_states(ga, oa)


def _anim_ops(levels, frames, scheme_merge="merge_random"):
    ops = []
    for lv in range(levels - 1, -1, -1):
        ops += [("style", lv), ("reset_adv",)]
        for f in range(frames):
            fwd = 1
            if f > 0:
                ops += [("flip",)]
            ops += [("frame", lv, f), ("params", dict(direction=_ffi.PT_FORWARD, frame_number=f))]
            if f == 0:
                ops += [("randomize",)] if lv == levels - 1 else [("upsample",), ("sync", 0)]
            else:
                ops += [("advect", fwd, _ffi.PT_ADVECT_UPDATE_RESIDUAL if scheme_merge == "merge_residual" else 0)]
                if lv < levels - 1:
                    ops += [("upsample",), ("sync", 0), (scheme_merge, fwd)]
            ops += [("sync", 0)]
            ops += [("propagate", j) for j in (8, 4, 2, 1, 1)] + [("scatter",)]
    return ops


@pytest.mark.parametrize("merge,hyst,tap,tt", [("merge_random", 0.3, 0, None), ("merge_residual", 0.3, 0, None),
                                               ("merge_random", 0.0, 0, None), ("merge_random", 0.3, 1024, None),
                                               ("merge_residual", 0.3, 200, None), ("merge_random", 0.3, 0, 0),
                                               ("merge_random", 0.0, 1024, 0)])
def test_animation_operator_sequence(merge, hyst, tap, tt):
    """Temporal terms against the oracle, with the per-level time tables
    (default) and without them (tt=0: every candidate gathers the advected
    offset's exemplar samples itself)."""
    st = W.style_pair(48, 48, seed=12, guides=True)
    frames = W.animation(40, 32, 3, seed=13)
    ov = dict(residual_window_size=2, levels=2, temporal_coherence_weight=0.5, hysteresis_weight=hyst)
    ga, oa = _both(st, frames, ov, _anim_ops(2, 3, merge), tap=tap, time_table=tt)
    assert_same_states(ga, oa)


@pytest.mark.parametrize("scheme", [_ffi.PT_S_FBMO, _ffi.PT_S_TCTF])
def test_time_tables_on_and_off_agree_bidirectional(scheme):
    """Fixup passes advect both neighbours in (bidirectional temporal terms,
    both tables live): the synthesis with and without the tables is the same
    bit for bit."""
    st = W.style_pair(64, 64, seed=14, guides=True)
    frames = W.animation(48, 40, 4, seed=15)
    kw = dict(residual_window_size=3, levels=3, temporal_coherence_weight=0.5)
    outs = []
    for tt in (81, 0):
        with Context.gpu(0) as g:
            g.set_tuning(_ffi.PT_TUNE_TIME_TABLE_MAX_TAPS, tt)
            outs.append(g.synthesize(g.default_params(**kw), frames, st, scheme=scheme, iterations=2, fixup_passes=1))
    _cmp_synth(outs[0], outs[1])


@pytest.mark.parametrize("early_exit,unit_parallel,tap", [(1, 1, 0), (0, 1, 0), (1, 0, 0), (0, 0, 0), (1, 1, 1024)])
@pytest.mark.parametrize("r", [2, 3])
def test_coherence_term_operator_sequence(r, early_exit, unit_parallel, tap):
    """The coherence term (against the current offsets of the tap's pixel)
    in every kernel shape, still schedule with orientation."""
    st = W.style_pair(40, 36, seed=15, guides=True)
    fr = W.still_frame(44, 34, seed=16, guides=True)
    ov = dict(residual_window_size=r, levels=3, coherence_weight=0.4, coherence_ratio_range=2.0,
              coherence_angle_range=0.25, orientation_enabled=1, dist_trans_weight=0.5)
    ga, oa = _both(st, [fr], ov, _still_ops(3), early_exit, unit_parallel, tap)
    assert_same_states(ga, oa)


@pytest.mark.parametrize("early_exit,unit_parallel,tap", [(1, 1, 0), (0, 0, 0), (1, 1, 1024)])
def test_lab_colour_space_operator_sequence(early_exit, unit_parallel, tap):
    """colorspace 1: every colour distance in L*a*b* (texSynth_kernel.cu:575-606)."""
    st = W.style_pair(40, 36, seed=19, guides=True)
    fr = W.still_frame(44, 34, seed=20, guides=True)
    ov = dict(residual_window_size=2, levels=3, colorspace=1, orientation_enabled=1, dist_trans_weight=0.5)
    ga, oa = _both(st, [fr], ov, _still_ops(3), early_exit, unit_parallel, tap)
    assert_same_states(ga, oa)


@pytest.mark.parametrize("early_exit,unit_parallel", [(1, 1), (0, 1), (0, 0)])
@pytest.mark.parametrize("r", [2, 3])
def test_extended_terms_in_unit_kernels(r, early_exit, unit_parallel):
    """Coherence, LAB colour space and object IDs together, evaluated by the
    unit-parallel propagate and scatter (their extended variants at a fixed
    radius, with and without early rejection) and by the per-pixel kernels,
    against the oracle (texSynth_kernel.cu:497-504,575-606,642-646)."""
    st, frs = _with_ids(W.style_pair(56, 48, seed=61, guides=True),
                        [W.still_frame(60, 44, seed=62, guides=True)], 63)
    ov = dict(residual_window_size=r, levels=3, coherence_weight=0.3, colorspace=1, style_obj_id_weight=0.6,
              orientation_enabled=1, dist_trans_weight=0.5)
    ga, oa = _both(st, frs, ov, _still_ops(3), early_exit, unit_parallel)
    assert_same_states(ga, oa)


@pytest.mark.parametrize("hyst,tt", [(0.3, None), (0.0, None), (0.3, 0), (0.0, 0)])
def test_lab_with_coherence_and_temporal_terms(hyst, tt):
    """LAB and coherence with the temporal terms: time tables stored in LAB
    (default) or samples gathered per candidate (tt=0); previous-frame colours
    staged converted by the unit kernels; hysteresis 0 lets the early
    rejection shapes run."""
    st = W.style_pair(48, 48, seed=21, guides=True)
    frames = W.animation(40, 32, 3, seed=22)
    ov = dict(residual_window_size=2, levels=2, temporal_coherence_weight=0.5, hysteresis_weight=hyst,
              coherence_weight=0.3, colorspace=1)
    ga, oa = _both(st, frames, ov, _anim_ops(2, 3, "merge_random"), time_table=tt)
    assert_same_states(ga, oa)


def test_extended_terms_bidirectional_passes():
    """FBMO fixup passes (both neighbours advected: bidirectional temporal
    terms, both time tables) with LAB, coherence and object IDs."""
    st, frames = _with_ids(W.style_pair(48, 48, seed=71, guides=True), W.animation(40, 32, 3, seed=72), 73)
    kw = dict(residual_window_size=3, levels=2, temporal_coherence_weight=0.5, coherence_weight=0.3, colorspace=1,
              style_obj_id_weight=0.6)
    with Context.gpu(0) as g, OracleContext() as o:
        _cmp_synth(g.synthesize(g.default_params(**kw), frames, st, scheme=_ffi.PT_S_FBMO, iterations=2),
                   o.synthesize(o.default_params(**kw), frames, st, scheme=_ffi.PT_S_FBMO, iterations=2))


@pytest.mark.parametrize("tap", [0, 1024])
def test_coherence_term_with_temporal_terms(tap):
    st = W.style_pair(48, 48, seed=17, guides=True)
    frames = W.animation(40, 32, 3, seed=18)
    ov = dict(residual_window_size=2, levels=2, temporal_coherence_weight=0.5, hysteresis_weight=0.3,
              coherence_weight=0.3)
    ga, oa = _both(st, frames, ov, _anim_ops(2, 3, "merge_random"), tap=tap)
    assert_same_states(ga, oa)


@pytest.mark.parametrize("window,ee", [(0, 0), (1, 0), (1, 1)])
@pytest.mark.parametrize("r", [2, 3])
def test_exemplar_window_refresh_and_propagate(r, window, ee):
    """The residual refresh and the short-jump propagate steps reading
    patches from an LDS window of the exemplar (PT_TUNE_EXEMPLAR_WINDOW,
    PT_TUNE_PROPAGATE_WINDOW_JUMP), with and without temporal terms and early
    rejection, against the oracle; exemplars from 96 down to 24 texels cover
    windows that clamp at the exemplar edges and levels too small for a
    window."""
    st = W.style_pair(96, 80, seed=33, guides=True)
    frames = W.animation(72, 64, 2, seed=34)
    ov = dict(residual_window_size=r, levels=3, temporal_coherence_weight=0.5, orientation_enabled=1)
    g = Context.gpu(0)
    g.set_tuning(_ffi.PT_TUNE_EXEMPLAR_WINDOW, window)
    g.set_tuning(_ffi.PT_TUNE_PROPAGATE_WINDOW_JUMP, 1 << 20)
    g.set_tuning(_ffi.PT_TUNE_PROPAGATE_WINDOW_MIN_PIXELS, window)
    g.set_tuning(_ffi.PT_TUNE_TAP_PARALLEL_MAX_PIXELS, 0)
    if ee:
        g.set_tuning(_ffi.PT_TUNE_EARLY_EXIT_MIN_PIXELS, 0)
    o = OracleContext()
    try:
        ga = Session(g, st, frames, g.default_params(**ov)).run(_anim_ops(3, 2))
        oa = Session(o, st, frames, o.default_params(**ov)).run(_anim_ops(3, 2))
    finally:
        g.close()
        o.close()
    assert_same_states(ga, oa)


@pytest.mark.parametrize("merge,early_exit,unit_parallel,tap", [("merge_random", 1, 1, 0), ("merge_residual", 1, 1, 0),
                                                               ("merge_random", 0, 0, 0), ("merge_random", 1, 1, 1024)])
def test_luminance_shift_operator_sequence(merge, early_exit, unit_parallel, tap):
    """lumShiftEnable: advection accumulates each offset's luminance shift,
    propagation and scatter reset it on the pixels they move, merges carry it,
    the canvas adds it (texSynth_kernel.cu:272, 1626-1631)."""
    st = W.style_pair(48, 48, seed=35, guides=True)
    frames = W.animation(40, 32, 3, seed=36)
    ov = dict(residual_window_size=2, levels=2, temporal_coherence_weight=0.5, hysteresis_weight=0.3,
              lum_shift_enable=1)
    ga, oa = _both(st, frames, ov, _anim_ops(2, 3, merge), early_exit, unit_parallel, tap)
    assert_same_states(ga, oa)
    assert any((state[0]["lum_y"] != 0).any() for _, state in ga)


def test_loaded_luminance_shifts():
    """Offsets loaded with shifts (pt_load_offsets) through propagate and
    scatter, against the oracle."""
    st = W.style_pair(40, 36, seed=37, guides=True)
    fr = W.still_frame(44, 34, seed=38, guides=True)
    rng = np.random.default_rng(39)
    outs = []
    for ctx in (Context.gpu(0), OracleContext()):
        try:
            s = Session(ctx, st, [fr], ctx.default_params(residual_window_size=2, levels=1))
            s.run([("style", 0), ("frame", 0, 0), ("randomize",)])
            off, _, _ = s.state(0)
            off["lum_y"] = rng.uniform(-0.2, 0.2, off.shape).astype(np.float32) if not outs else outs[0][0]["lum_y"]
            ctx.load_offsets(0, off)
            ctx.sync_buffers(0, _ffi.PT_DONT_COPY_OFFSETS)
            res = [s.state(0)]
            res += [st_ for _, st_ in s.run([("propagate", 4), ("propagate", 1), ("scatter",), ("propagate", 1)])]
            outs.append((off, res))
        finally:
            ctx.close()
    for a, b in zip(outs[0][1], outs[1][1]):
        for x, y in zip(a, b):
            assert first_difference(x, y) is None


@pytest.mark.parametrize("styles", [2, 4])
def test_several_styles_operator_sequence(styles):
    """numStyles > 1 without keyframes: offsets stay in style 0, taps wrap at
    style height x numStyles (clamping to the exemplar's last row below it)."""
    st = W.style_pair(40, 36, seed=42, guides=True)
    fr = W.still_frame(44, 34, seed=43, guides=True)
    ov = dict(residual_window_size=3, levels=3, num_styles=styles, orientation_enabled=1, dist_trans_weight=0.5)
    ga, oa = _both(st, [fr], ov, _still_ops(3))
    assert_same_states(ga, oa)
    ga1, _ = _both(st, [fr], dict(ov, num_styles=1), _still_ops(3)[:5])
    assert ga[4][1][2].tobytes() != ga1[4][1][2].tobytes()  # the wrap changed the residuals


def test_synthesize_several_styles_bit_exact():
    st = W.style_pair(64, 64, seed=44, guides=True)
    frames = W.animation(48, 40, 3, seed=45)
    with Context.gpu(0) as g, OracleContext() as o:
        kw = dict(residual_window_size=3, levels=3, temporal_coherence_weight=0.5, num_styles=3)
        _cmp_synth(g.synthesize(g.default_params(**kw), frames, st, iterations=2),
                   o.synthesize(o.default_params(**kw), frames, st, iterations=2))


def _with_ids(st, frames, seed):
    """Attach object-ID images (Voronoi cells) to a style and its frames."""
    sh, sw = st["base"].shape[:2]
    st = dict(st, object_ids=W.object_ids(seed, 30, sw, sh))
    out = []
    for i, f in enumerate(frames):
        h, w = f["color"].shape[:2]
        out.append(dict(f, object_id=W.object_ids(seed, 31 + i, w, h)))
    return st, out


def test_object_id_pyramid_bit_exact():
    """downsampleVote on the GPU equals the oracle's, odd sizes included."""
    with Context.gpu(0) as g, OracleContext() as o:
        for (w, h), seed in (((45, 27), 1), ((64, 64), 2)):
            ids = W.object_ids(seed, 3, w, h, regions=9)
            ids[::3, ::2] = 5  # ties and majorities of every kind
            pg, po = g.pyramid(_ffi.PT_ELEM_I32, ids), o.pyramid(_ffi.PT_ELEM_I32, ids)
            assert pg.levels == po.levels
            for l in range(pg.levels):
                assert pg.download(l).tobytes() == po.download(l).tobytes(), ((w, h), l)


@pytest.mark.parametrize("early_exit,unit_parallel,tap", [(1, 1, 0), (0, 0, 0), (1, 1, 1024)])
def test_object_id_term_operator_sequence(early_exit, unit_parallel, tap):
    """styleObjIdWeight > 0: the object-ID term at the centre tap
    (texSynth_kernel.cu:497-504, 637-638), still schedule."""
    st, frs = _with_ids(W.style_pair(40, 36, seed=23, guides=True), [W.still_frame(44, 34, seed=24, guides=True)],
                        25)
    ov = dict(residual_window_size=2, levels=3, style_obj_id_weight=0.6, orientation_enabled=1, dist_trans_weight=0.5)
    ga, oa = _both(st, frs, ov, _still_ops(3), early_exit, unit_parallel, tap)
    assert_same_states(ga, oa)


def test_object_ids_with_temporal_and_coherence_terms():
    st, frames = _with_ids(W.style_pair(48, 48, seed=26, guides=True), W.animation(40, 32, 3, seed=27), 28)
    ov = dict(residual_window_size=2, levels=2, temporal_coherence_weight=0.5, hysteresis_weight=0.3,
              coherence_weight=0.3, style_obj_id_weight=0.5)
    ga, oa = _both(st, frames, ov, _anim_ops(2, 3, "merge_random"))
    assert_same_states(ga, oa)


def test_kernel_variants_agree_at_scale():
    """Early rejection, unit-parallel and tap-parallel search and the XCD
    tile mapping on and off, on a 256^2 still with guides: offsets, canvas
    and residual identical."""
    st = W.style_pair(256, 256, seed=51, guides=True)
    fr = W.still_frame(256, 256, seed=52, guides=True)
    outs = []
    for ee, up, tap, xcd in ((1, 1, 128 * 128, 1), (0, 0, 0, 1), (1, 0, 0, 0), (0, 1, 0, 1), (1, 1, 1 << 30, 1),
                             (1, 1, 128 * 128, 0)):
        with Context.gpu(0) as g:
            g.set_tuning(_ffi.PT_TUNE_EARLY_EXIT, ee)
            g.set_tuning(_ffi.PT_TUNE_UNIT_PARALLEL, up)
            g.set_tuning(_ffi.PT_TUNE_EARLY_EXIT_MIN_PIXELS, 0)
            g.set_tuning(_ffi.PT_TUNE_TAP_PARALLEL_MAX_PIXELS, tap)
            g.set_tuning(_ffi.PT_TUNE_XCD_TILES, xcd)
            p = g.default_params(residual_window_size=3, levels=4, dist_trans_weight=0.5)
            outs.append(g.synthesize(p, [fr], st, iterations=4))
    for other in outs[1:]:
        _cmp_synth(outs[0], other)


def test_headline_config_coarse_levels_bit_exact():
    """The benchmark still's configuration (1024^2 inputs, 7x7 patches, 5
    levels, 8 iterations, guides), default kernel policy: every state of
    levels 4 to 1 (64^2 to 512^2) equals the oracle's; level 0 is checked by
    the full-size consistency test below."""
    st = W.style_pair(1024, 1024, seed=2, guides=True)
    fr = W.still_frame(1024, 1024, seed=2, guides=True)
    kw = dict(residual_window_size=3, levels=5, dist_trans_weight=0.5, orientation_enabled=1)
    ops = _still_ops(5, iterations=8)
    ops = ops[:ops.index(("style", 0))]  # levels 4 to 1
    g, o = Context.gpu(0), OracleContext()
    try:
        ga = Session(g, st, [fr], g.default_params(**kw)).run(ops)
        oa = Session(o, st, [fr], o.default_params(**kw)).run(ops)
    finally:
        g.close()
        o.close()
    assert_same_states(ga, oa)


def test_headline_still_full_size_consistent_with_oracle():
    """The benchmark still at its full size (1024^2 frame and exemplar pair,
    7x7 patches, 5 levels, 8 iterations, orientation and distance-transform
    guides): the oracle, given the GPU's final offsets, recomputes the canvas
    and every pixel's residual bit for bit; a second GPU run reproduces the
    first; offsets stay inside the exemplar."""
    st = W.style_pair(1024, 1024, seed=2, guides=True)
    fr = W.still_frame(1024, 1024, seed=2, guides=True)
    kw = dict(residual_window_size=3, levels=5, dist_trans_weight=0.5, orientation_enabled=1)
    with Context.gpu(0) as g:
        p = g.default_params(**kw)
        a = g.synthesize(p, [fr], st, iterations=8)
        b = g.synthesize(p, [fr], st, iterations=8)
    _cmp_synth(a, b)
    off, canvas, res = (x[0] for x in a)
    for f in ("x", "y"):
        assert off[f].min() >= 0.0 and off[f].max() <= 1025.0
    assert np.isfinite(res).all() and (res >= 0).all()
    with OracleContext() as o:
        s = Session(o, st, [fr], o.default_params(**kw))
        s.run([("style", 0), ("frame", 0, 0)])
        o.load_offsets(0, off)
        o.sync_buffers(0, _ffi.PT_DONT_COPY_OFFSETS)
        _, oc, orr = s.state(0)
    assert first_difference(canvas, oc) is None, "canvas"
    assert first_difference(res, orr) is None, "residual"


def _cmp_synth(g_out, o_out):
    for label, a, b in zip(("offsets", "canvas", "residual"), g_out, o_out):
        assert first_difference(a, b) is None, label


def test_synthesize_still_bit_exact():
    st = W.style_pair(64, 64, seed=21, guides=True)
    fr = W.still_frame(64, 48, seed=22, guides=True)
    with Context.gpu(0) as g, OracleContext() as o:
        kw = dict(residual_window_size=3, levels=3, dist_trans_weight=0.5)
        _cmp_synth(g.synthesize(g.default_params(**kw), [fr], st, iterations=3),
                   o.synthesize(o.default_params(**kw), [fr], st, iterations=3))


@pytest.mark.parametrize("scheme", [_ffi.PT_S_FMBM_RM, _ffi.PT_S_FMBM])
def test_synthesize_animation_bit_exact(scheme):
    st = W.style_pair(64, 64, seed=23, guides=True)
    frames = W.animation(48, 40, 4, seed=24)
    with Context.gpu(0) as g, OracleContext() as o:
        kw = dict(residual_window_size=3, levels=3, temporal_coherence_weight=0.5)
        _cmp_synth(g.synthesize(g.default_params(**kw), frames, st, scheme=scheme, iterations=2, fixup_passes=1),
                   o.synthesize(o.default_params(**kw), frames, st, scheme=scheme, iterations=2, fixup_passes=1))


@pytest.mark.parametrize("scheme", [_ffi.PT_S_FMBM_RM, _ffi.PT_S_TCTF])
def test_synthesize_object_ids_bit_exact(scheme):
    """Object-ID pyramids built from the guides (vote downsample) and
    uploaded beside the orientation and distance transform each level."""
    st, frames = _with_ids(W.style_pair(64, 64, seed=29, guides=True), W.animation(48, 40, 3, seed=30), 31)
    with Context.gpu(0) as g, OracleContext() as o:
        kw = dict(residual_window_size=3, levels=3, temporal_coherence_weight=0.5, style_obj_id_weight=0.7)
        _cmp_synth(g.synthesize(g.default_params(**kw), frames, st, scheme=scheme, iterations=2, fixup_passes=1),
                   o.synthesize(o.default_params(**kw), frames, st, scheme=scheme, iterations=2, fixup_passes=1))


@pytest.mark.parametrize("scheme", [_ffi.PT_S_FMBM_RM, _ffi.PT_S_FMBM, _ffi.PT_S_FBMO])
def test_synthesize_luminance_shift_bit_exact(scheme):
    st = W.style_pair(64, 64, seed=40, guides=True)
    frames = W.animation(48, 40, 3, seed=41)
    with Context.gpu(0) as g, OracleContext() as o:
        kw = dict(residual_window_size=3, levels=3, temporal_coherence_weight=0.5, lum_shift_enable=1)
        a = g.synthesize(g.default_params(**kw), frames, st, scheme=scheme, iterations=2, fixup_passes=1)
        _cmp_synth(a, o.synthesize(o.default_params(**kw), frames, st, scheme=scheme, iterations=2, fixup_passes=1))
    assert (a[0]["lum_y"] != 0).any()


@pytest.mark.parametrize("scheme", [_ffi.PT_S_FMBM_RM, _ffi.PT_S_FMBM, _ffi.PT_S_FBMO])
def test_synthesize_backward_first_pass_bit_exact(scheme):
    """firstPassDirection 1: pass p runs forwards iff p % 2 == 1."""
    st = W.style_pair(64, 64, seed=46, guides=True)
    frames = W.animation(48, 40, 3, seed=47)
    with Context.gpu(0) as g, OracleContext() as o:
        kw = dict(residual_window_size=3, levels=3, temporal_coherence_weight=0.5, first_pass_direction=1)
        _cmp_synth(g.synthesize(g.default_params(**kw), frames, st, scheme=scheme, iterations=2, fixup_passes=1),
                   o.synthesize(o.default_params(**kw), frames, st, scheme=scheme, iterations=2, fixup_passes=1))


@pytest.mark.parametrize("scheme,max_level", [(_ffi.PT_S_FMBM_RM, 1), (_ffi.PT_S_FMBM, 2), (_ffi.PT_S_FMBM, 1)])
def test_synthesize_max_level_bit_exact(scheme, max_level):
    """maxLevel > 0: synthesis stops at that level, then every frame's result
    is upsampled level by level to the full size (renderAnimationFMBM's
    tail, synthesisProcessor.cpp:863-870); other schemes reject it."""
    st = W.style_pair(64, 64, seed=48, guides=True)
    frames = W.animation(48, 40, 3, seed=49)
    with Context.gpu(0) as g, OracleContext() as o:
        kw = dict(residual_window_size=3, levels=3, temporal_coherence_weight=0.5, max_level=max_level)
        _cmp_synth(g.synthesize(g.default_params(**kw), frames, st, scheme=scheme, iterations=2, fixup_passes=1),
                   o.synthesize(o.default_params(**kw), frames, st, scheme=scheme, iterations=2, fixup_passes=1))
        for other in (_ffi.PT_S_IND, _ffi.PT_S_FBMO, _ffi.PT_S_TCTF):
            with pytest.raises(Exception):
                g.synthesize(g.default_params(**kw), frames, st, scheme=other, iterations=2, fixup_passes=1)


def test_synthesize_histogram_slope_bit_exact():
    """Offsets-histogram slopes are accepted where the reference's term is 0
    (an all-zero histogram, thresholds not negative) and change nothing; a
    negative layer-0 threshold with a slope is refused."""
    st = W.style_pair(64, 64, seed=50, guides=True)
    frames = W.animation(48, 40, 2, seed=51)
    with Context.gpu(0) as g, OracleContext() as o:
        kw = dict(residual_window_size=3, levels=3, temporal_coherence_weight=0.5)
        pg, po, p0 = g.default_params(**kw), o.default_params(**kw), g.default_params(**kw)
        for p in (pg, po):
            p.offsets_histogram_slope[0] = 2.0
            p.offsets_histogram_threshold[0] = 0.3
        got = g.synthesize(pg, frames, st, scheme=_ffi.PT_S_FMBM, iterations=2, fixup_passes=1)
        _cmp_synth(got, o.synthesize(po, frames, st, scheme=_ffi.PT_S_FMBM, iterations=2, fixup_passes=1))
        _cmp_synth(got, g.synthesize(p0, frames, st, scheme=_ffi.PT_S_FMBM, iterations=2, fixup_passes=1))
        pg.offsets_histogram_threshold[0] = -0.5
        with pytest.raises(Exception):
            g.synthesize(pg, frames, st, scheme=_ffi.PT_S_FMBM, iterations=2, fixup_passes=1)


@pytest.mark.parametrize("fixup", [0, 1])
def test_synthesize_ind_bit_exact(fixup):
    """renderAnimationIND: every frame on its own, levels inside frames."""
    st = W.style_pair(64, 64, seed=65, guides=True)
    frames = W.animation(48, 40, 3, seed=66)
    with Context.gpu(0) as g, OracleContext() as o:
        kw = dict(residual_window_size=3, levels=3, temporal_coherence_weight=0.5, hysteresis_weight=0.2)
        _cmp_synth(g.synthesize(g.default_params(**kw), frames, st, scheme=_ffi.PT_S_IND, iterations=2,
                                fixup_passes=fixup),
                   o.synthesize(o.default_params(**kw), frames, st, scheme=_ffi.PT_S_IND, iterations=2,
                                fixup_passes=fixup))


@pytest.mark.parametrize("fixup", [0, 1])
def test_synthesize_fbmo_bit_exact(fixup):
    """renderAnimationFBMO: forward and backward passes, their merge, and
    fixup passes with both neighbours advected in (bidirectional temporal
    terms)."""
    st = W.style_pair(64, 64, seed=67, guides=True)
    frames = W.animation(48, 40, 3, seed=68)
    with Context.gpu(0) as g, OracleContext() as o:
        kw = dict(residual_window_size=3, levels=3, temporal_coherence_weight=0.5, hysteresis_weight=0.2)
        _cmp_synth(g.synthesize(g.default_params(**kw), frames, st, scheme=_ffi.PT_S_FBMO, iterations=2,
                                fixup_passes=fixup),
                   o.synthesize(o.default_params(**kw), frames, st, scheme=_ffi.PT_S_FBMO, iterations=2,
                                fixup_passes=fixup))


@pytest.mark.parametrize("fixup", [0, 1])
def test_synthesize_tctf_bit_exact(fixup):
    """renderAnimationTCTF: a consecutive pass, then refine passes coarse to
    fine with both neighbours advected in."""
    st = W.style_pair(64, 64, seed=71, guides=True)
    frames = W.animation(48, 40, 3, seed=72)
    with Context.gpu(0) as g, OracleContext() as o:
        kw = dict(residual_window_size=3, levels=3, temporal_coherence_weight=0.5, hysteresis_weight=0.2)
        _cmp_synth(g.synthesize(g.default_params(**kw), frames, st, scheme=_ffi.PT_S_TCTF, iterations=2,
                                fixup_passes=fixup),
                   o.synthesize(o.default_params(**kw), frames, st, scheme=_ffi.PT_S_TCTF, iterations=2,
                                fixup_passes=fixup))


def test_config1_still_256_bit_exact():
    """The CPU reference workload: 256^2 exemplar pair and target, 3 levels,
    5x5 patches, 4 iterations, orientation off, no temporal term."""
    st = W.style_pair(256, 256, seed=1)
    fr = W.still_frame(256, 256, seed=1)
    with Context.gpu(0) as g, OracleContext() as o:
        kw = dict(residual_window_size=2, levels=3, orientation_enabled=0, dist_trans_weight=0.0)
        _cmp_synth(g.synthesize(g.default_params(**kw), [fr], st, iterations=4),
                   o.synthesize(o.default_params(**kw), [fr], st, iterations=4))


def test_repeat_calls_identical():
    """Reused device allocations give the same result on a second call."""
    st = W.style_pair(48, 48, seed=31)
    fr = W.still_frame(40, 40, seed=32)
    with Context.gpu(0) as g:
        p = g.default_params(residual_window_size=2, levels=2)
        a = g.synthesize(p, [fr], st, iterations=2)
        b = g.synthesize(p, [fr], st, iterations=2)
    for x, y in zip(a, b):
        assert x.tobytes() == y.tobytes()


@pytest.mark.parametrize("anim", [False, True])
def test_incremental_residual_matches_full_refresh(anim):
    """Skipping pixels whose patch window saw no offset change reproduces the
    full refresh bit for bit (still and animation with temporal terms)."""
    if anim:
        st = W.style_pair(96, 80, seed=91, guides=True)
        frames = W.animation(128, 96, 3, seed=92)
        kw = dict(residual_window_size=3, levels=3, temporal_coherence_weight=0.5)
    else:
        st = W.style_pair(192, 160, seed=93, guides=True)
        frames = [W.still_frame(256, 224, seed=94, guides=True)]
        kw = dict(residual_window_size=3, levels=4, dist_trans_weight=0.5)
    outs = []
    for inc in (1, 0):
        with Context.gpu(0) as g:
            g.set_tuning(_ffi.PT_TUNE_INCREMENTAL_RESIDUAL, inc)
            outs.append(g.synthesize(g.default_params(**kw), frames, st, iterations=3, fixup_passes=1))
    for a, b in zip(*outs):
        assert a.tobytes() == b.tobytes()


# ---------------------------------------------------------------- culling
def _culling_inputs(w, h, sw, sh, frames=1, seed=60):
    """Distance transforms in pixel units straddling the culling thresholds:
    image 30..70 around the cull 50, exemplar 40..80 around 1.25 * 50, and
    (animations) previous frames around cull - 8 = 42."""
    st = W.style_pair(sw, sh, seed=seed, guides=True)
    st["dist_trans"] = W.culling_dist_trans(seed, 1, sw, sh, lo=40.0, hi=80.0)
    if frames == 1:
        fr = [W.still_frame(w, h, seed=seed + 1, guides=True)]
    else:
        fr = W.animation(w, h, frames, seed=seed + 1)
    for k, f in enumerate(fr):
        f["dist_trans"] = W.culling_dist_trans(seed, 10 + k, w, h)
    return st, fr


@pytest.mark.parametrize("early_exit,unit_parallel,tap", [(1, 1, 0), (0, 1, 0), (1, 0, 0), (0, 0, 0), (1, 1, 1024),
                                                         (1, 1, 300)])
@pytest.mark.parametrize("r", [2, 3])
def test_culling_still_operator_sequence(r, early_exit, unit_parallel, tap):
    """Culled pixels (zero offset, residual 0, no scatter write) and culled
    exemplar texels (candidates rejected) in every kernel shape."""
    st, fr = _culling_inputs(44, 34, 40, 36)
    ov = dict(residual_window_size=r, levels=3, dist_trans_weight=0.5)
    ga, oa = _both(st, fr, ov, _still_ops(3), early_exit, unit_parallel, tap)
    assert_same_states(ga, oa)
    off, _, res = ga[-1][1]
    culled = W.culling_dist_trans(60, 10, 44, 34) > 50
    assert culled.any() and (~culled).any()
    assert np.all(off["x"][culled] == 0) and np.all(res[culled] == 0)


@pytest.mark.parametrize("merge,tap", [("merge_random", 0), ("merge_residual", 0), ("merge_random", 1024)])
def test_culling_animation_operator_sequence(merge, tap):
    """Advection with previous-frame culling (cull - 8) and the culled
    current pixels, temporal terms on."""
    st, frames = _culling_inputs(40, 32, 48, 48, frames=3)
    ov = dict(residual_window_size=2, levels=2, temporal_coherence_weight=0.5, hysteresis_weight=0.3,
              dist_trans_weight=0.5)
    ga, oa = _both(st, frames, ov, _anim_ops(2, 3, merge), tap=tap)
    assert_same_states(ga, oa)


@pytest.mark.parametrize("scheme", [_ffi.PT_S_FMBM_RM, _ffi.PT_S_FMBM])
def test_culling_synthesize_animation_bit_exact(scheme):
    st, frames = _culling_inputs(48, 40, 64, 64, frames=3, seed=70)
    with Context.gpu(0) as g, OracleContext() as o:
        kw = dict(residual_window_size=3, levels=3, temporal_coherence_weight=0.5, dist_trans_weight=0.5)
        a = g.synthesize(g.default_params(**kw), frames, st, scheme=scheme, iterations=2, fixup_passes=1)
        _cmp_synth(a, o.synthesize(o.default_params(**kw), frames, st, scheme=scheme, iterations=2, fixup_passes=1))
    assert (a[2] == 0).any()
