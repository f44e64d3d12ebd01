"""GPU parity of keyframe interpolation: the HIP library against the CPU
oracle, bit for bit, through the C ABI.  After every operator the base
offsets (style layer included), canvas and residual, both advected sets
and the ribbon field are compared as raw bytes.

Covered: the *_keys propagate and scatter kernels (walks through the
cumulative key-to-key fields, the keymask of a keyframe index, samples
shared by the keys), random initialisation drawing the style per pixel,
advection with ADVECT_CHECK_FRAME_TO_KEY, mergeRandomly's ribbon clause,
mergeAdvectedBuffers with both fallbacks, offsetsFromFrameToKeyRibbon,
ribbon-field accumulation with its canvas lookup, ribbon pyramids, and
layered exemplars without keyframes (offsets loaded with layers)."""
import numpy as np
import pytest

from painttween_amd import Context, _ffi
from painttween_amd import workloads as W
from tests.harness import first_difference
from tests.keyframes import KeySession, RIBBON, assert_same_key_states, key_params, key_styles, ribbon_texels
from tests.oracle_ctx import OracleContext

pytestmark = pytest.mark.gpu

CHECK = _ffi.PT_ADVECT_CHECK_FRAME_TO_KEY
UPDATE = _ffi.PT_ADVECT_UPDATE_RESIDUAL


def _both(w, h, nk, nframes, overrides, ops, seed=5):
    styles = key_styles(w, h, nk, seed)
    frames = W.animation(w, h, nframes, seed=seed)
    g, o = Context.gpu(0), OracleContext()
    try:
        sg = KeySession(g, styles, frames, key_params(g, nk, **overrides), seed=seed)
        so = KeySession(o, styles, frames, key_params(o, nk, **overrides), seed=seed)
        return sg.run(ops), so.run(ops)
    finally:
        g.close()
        o.close()


def _search(level, keys, iterations=1):
    ops = []
    for _ in range(iterations):
        ops += [("propagate", j, keys[i % len(keys)]) for i, j in enumerate((8, 4, 2, 1))]
        ops += [("scatter", keys[0]), ("scatter", -1)]
    return ops


@pytest.mark.parametrize("r", [2, 3])
def test_keyed_search_two_levels(r):
    """Three keys, coarse level from the per-pixel style draw, then the
    fine level from the upsampled field; propagations with and without a
    keyframe index, scatters with the keys' shared samples."""
    ops = [("styles", 1), ("frame", 1, 0), ("randomize",)] + _search(1, [1, -1, 0], 2)
    ops += [("styles", 0), ("frame", 0, 0), ("upsample",), ("sync", 0)] + _search(0, [-1, 2])
    ga, oa = _both(56, 44, 3, 1, dict(residual_window_size=r), ops)
    assert_same_key_states(ga, oa)
    off = ga[-1][1][0]
    assert set(np.unique(off["layer"])) <= {0, 1, 2} and len(np.unique(off["layer"])) > 1


def test_keyed_animation_frame_to_key_advection_and_merges():
    """Two frames: the second advected forwards with the frame-to-key check
    (time steps 1 and 3), merged randomly and by residual, the advected
    pair merged with both fallbacks, and offsets taken from the ribbons in
    both directions; temporal terms with hysteresis on."""
    ov = dict(temporal_coherence_weight=0.5, hysteresis_weight=0.3, direction=_ffi.PT_BIDIRECTIONAL,
              advection_jitter=0.25)
    ops = [("styles", 0), ("frame", 0, 0), ("randomize",)] + _search(0, [0, -1])
    ops += [("flip",), ("frame", 0, 1), ("advect", True, 1, CHECK | UPDATE), ("merge_random", True, 0.6)]
    ops += _search(0, [-1, 1])
    ops += [("advect", False, 3, CHECK), ("merge_residual", False), ("merge_advected", _ffi.PT_MERGE_FALLBACK_RANDOM)]
    ops += _search(0, [2])
    ops += [("advect", True, 3, CHECK | UPDATE), ("merge_advected", _ffi.PT_MERGE_FALLBACK_LAST_PASS),
            ("merge_random", False, 0.5), ("ftk_offsets", True), ("residual",), ("ftk_offsets", False),
            ("propagate", 1, -1), ("scatter", -1)]
    ga, oa = _both(48, 40, 3, 2, ov, ops)
    assert_same_key_states(ga, oa)


def test_ribbon_field_accumulation():
    """The ribbon field advected forwards and backwards, overwriting and
    keeping the shorter ribbons, with the canvas lookup into the forward
    advected output; at the coarse level and at level 0."""
    ops = [("styles", 1), ("frame", 1, 0), ("randomize",), ("propagate", 2, -1), ("load_ribbons", 9, 1),
           ("flip",), ("frame", 1, 1), ("accumulate", True, 1, True), ("accumulate", False, 2, False),
           ("accumulate", True, 3, False)]
    ops += [("styles", 0), ("frame", 0, 1), ("randomize",), ("load_ribbons", 9, 2), ("accumulate", True, 1, False),
            ("accumulate", False, 1, True), ("advect", True, 2, CHECK)]
    ga, oa = _both(48, 40, 3, 2, {}, ops)
    assert_same_key_states(ga, oa)


def test_layered_exemplar_without_keys():
    """Several exemplar layers without keyframe interpolation: offsets
    loaded with layers 0..2 sample their own layer in the residual, the
    canvas and every search; scatter's samples and averaged candidate stay
    in layer 0, propagation and upsampling carry the neighbour's layer."""
    w, h = 48, 40
    rng = np.random.default_rng(3)
    off = np.zeros((h >> 1, w >> 1), _ffi.PATCHXF_DTYPE)
    off["x"] = rng.uniform(0, (w >> 1) - 1, off.shape)
    off["y"] = rng.uniform(0, (h >> 1) - 1, off.shape)
    off["layer"] = rng.integers(0, 3, off.shape)
    off["scale_u"] = off["scale_v"] = 1.0
    ops = [("styles", 1), ("frame", 1, 0, False), ("load_offsets", off), ("sync", _ffi.PT_DONT_COPY_OFFSETS),
           ("residual",), ("propagate", 2, -1),
           ("scatter", -1), ("propagate", 1, -1), ("styles", 0), ("frame", 0, 0, False), ("upsample",), ("sync", 0),
           ("scatter", -1), ("propagate", 4, -1)]
    ga, oa = _both(w, h, 3, 1, dict(interpolate_key_frame=0, first_key_index=-1, last_key_index=-1,
                                    temporal_coherence_weight=0.0), ops)
    assert_same_key_states(ga, oa)


def test_ribbon_pyramid():
    t = ribbon_texels(40, 36, 3, 4, 7)
    with Context.gpu(0) as g, OracleContext() as o:
        pg, po = g.pyramid(_ffi.PT_ELEM_F32X4, t, RIBBON), o.pyramid(_ffi.PT_ELEM_F32X4, t, RIBBON)
        assert pg.levels == po.levels
        for l in range(pg.levels):
            assert first_difference(pg.download(l), po.download(l)) is None, l
