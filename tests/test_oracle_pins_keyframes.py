"""Pins of the oracle's keyframe-interpolation operators by closed forms
and invariants, chosen so that a plausible mistake in the restatement (a
walk that reads the wrong field layer or swaps the x and y channels, a
stop rule on the wrong channel, a bilinear read where the reference reads
a texel, a missing divide, a swapped direction) fails here:

  propagate_kernel_keys / scatter_kernel_keys  texSynth_kernel.cu:723-839, 984-1110
  keymask                                      :887-892, 1137-1142
  layered sampling (PYRAMID_FETCH_LAYER)       cudaImagePyramid.h:108
  advectKernel ADVECT_CHECK_FRAME_TO_KEY       :1643-1668
  mergeRandomlyKernel ribbon clause            :1891-1936
  mergeAdvectedKernel                          :1938-1976
  copyFrameToKeyRibbonKernel                   :2061-2097
  accumulateAdvectionKernel + canvas lookup    :1728-1826, 290-305
  initializeRibbonPyramid                      imagePyramid.cpp:211-235
  createRandomPyramid with keys                synthesisFrame.cpp:102-143

Images are small and the expectations are stated from the reference's
definitions, not by replaying the oracle's code."""
import numpy as np
import pytest

from painttween_amd import _ffi
from painttween_amd import workloads as W
from tests.oracle_ctx import OracleContext

E1, E2, E3, E4 = _ffi.PT_ELEM_F32, _ffi.PT_ELEM_F32X2, _ffi.PT_ELEM_F32X3, _ffi.PT_ELEM_F32X4
CONST, WITH_LEVELS, RIBBON = (_ffi.PT_DOWNSAMPLE_SCALE_CONSTANT, _ffi.PT_DOWNSAMPLE_SCALE_WITH_LEVELS,
                              _ffi.PT_DOWNSAMPLE_RIBBON)
CHECK = _ffi.PT_ADVECT_CHECK_FRAME_TO_KEY


class Rig:
    """An oracle context at level 0 with nk exemplar layers, one or two
    frames and optional cumulative fields; guides neutral unless given
    (orientation off, no culling, scale 1)."""

    def __init__(self, o, frames, layers, keys=True, r=1, cumul=None, **kw):
        self.o = o
        h, w = frames[0]["color"].shape[:2]
        sh, sw = layers[0]["base"].shape[:2]
        nk = len(layers)
        self.w, self.h, self.sw, self.sh, self.nk = w, h, sw, sh, nk
        prm = dict(residual_window_size=r, levels=1, num_styles=nk, orientation_enabled=0,
                   canvas_output_weight=0.0, dist_trans_weight=0.0, transparency_ok=1,
                   time_derivative_input_weight=0.0, time_derivative_output_weight=0.0,
                   temporal_coherence_weight=0.0)
        if keys:
            prm.update(interpolate_key_frame=1, first_key_index=0, last_key_index=nk - 1)
        prm.update(kw)
        self.params = o.default_params(**prm)
        o.setup(self.params)
        o.initialize(w, h, sw, sh)
        o.initialize_exemplars(sw, sh, nk)
        self.keep = []
        for k, L in enumerate(layers):
            for name, arr, elem in (("EXEMPLAR_BASE", L["base"], E4), ("EXEMPLAR_OUTPUT", L["output"], E4)):
                p = o.pyramid(elem, arr)
                o.upload_image(0, name, p, k)
                self.keep.append(p)
            for name in ("EXEMPLAR_ORIENTATION", "EXEMPLAR_DIST_TRANS"):
                p = o.pyramid(E1, None, shape=(sw, sh), default=0.0)
                o.upload_image(0, name, p, k)
                self.keep.append(p)
        if cumul is not None:
            o.initialize_cumulative_fields(sw, sh, nk)
            for k in range(nk):
                for name, arr in (("CUMUL_FIELD_B", cumul[0][k]), ("CUMUL_FIELD_F", cumul[1][k])):
                    p = o.pyramid(E4, arr)
                    o.upload_image(0, name, p, k)
                    self.keep.append(p)
        self.frames = frames
        self.load_frame(0)

    def pyr(self, elem, arr, mode=CONST, default=0.0):
        p = self.o.pyramid(elem, arr, mode, default=default, shape=(self.w, self.h))
        self.keep.append(p)
        return p

    def load_frame(self, f, random=True):
        o, fr = self.o, self.frames[f]
        col = self.pyr(E4, fr["color"])
        ori = self.pyr(E1, None)
        o.upload_image(0, "INPUT_COLOR", col)
        o.upload_image(0, "INPUT_ORIENTATION", ori)
        o.upload_image(0, "INPUT_SCALE", self.pyr(E2, None, default=1.0))
        o.upload_image(0, "INPUT_DIST_TRANS", self.pyr(E1, None, default=1.0))
        o.upload_image(0, "INPUT_VEL_B", self.pyr(E2, fr.get("vel_b"), WITH_LEVELS))
        o.upload_image(0, "INPUT_VEL_F", self.pyr(E2, fr.get("vel_f"), WITH_LEVELS))
        for name in ("FRAME_TO_KEY_RIBBON_B", "FRAME_TO_KEY_RIBBON_F"):
            if name in fr:
                o.upload_image(0, name, self.pyr(E4, fr[name]))
        if random:
            eo = self.o.pyramid(E4, np.zeros((self.sh, self.sw, 4), np.float32) + 1.0)
            er = self.o.pyramid(E1, None, shape=(self.sw, self.sh))
            self.keep += [eo, er]
            o.create_random_pyramid(0, f, col, ori, eo, er)

    def offsets(self, layer=0, dx=0.0, dy=0.0):
        """Offsets at each pixel's own position + (dx, dy) in `layer`."""
        off = np.zeros((self.h, self.w), _ffi.PATCHXF_DTYPE)
        yy, xx = np.mgrid[0:self.h, 0:self.w]
        off["x"], off["y"] = xx + dx, yy + dy
        off["layer"] = layer
        off["scale_u"] = off["scale_v"] = 1.0
        return off

    def load(self, off):
        """Base offsets, then the canvas and residual for them."""
        self.o.load_offsets(0, off)
        self.o.sync_buffers(0, _ffi.PT_DONT_COPY_OFFSETS)

    def settle(self):
        """Working offsets = base offsets (a mergeRandomly of weight 0 with
        no frame-to-key ribbons copies the base), as after a sync."""
        self.o.merge_randomly(0, True, 0.0)

    def out(self):
        return self.o.get_output(0, self.w, self.h)


def _shift(img, dx, dy):
    """img shifted by (dx, dy): out(x, y) = img((x - dx) mod w, (y - dy) mod h)."""
    return np.roll(np.roll(img, dy, axis=0), dx, axis=1)


def _const_field(sw, sh, ts, dy, dx, layer):
    f = np.zeros((sh, sw, 4), np.float32)
    f[..., 0], f[..., 1], f[..., 2], f[..., 3] = ts, dy, dx, layer
    return f


def _walk_rig(o, forward, step_ts=0.0, r=1, seed=11):
    """Image P; exemplar layers 0 and 1 are unrelated noise, the far key's
    layer holds P displaced by two field steps, so a walk that follows the
    constant field (dx, dy) per key reaches an exact match."""
    w, h = 40, 32
    dx, dy = (3, 2) if forward else (-3, 2)
    P = W.sin_noise(seed, 0, w, h)
    noise = [W.sin_noise(seed, 1 + k, w, h) for k in range(3)]
    far = 2 if forward else 0
    layers = []
    for k in range(3):
        b = _shift(P, 2 * dx, 2 * dy) if k == far else noise[k]
        layers.append({"base": b, "output": b})
    zero = [_const_field(w, h, 0.0, 0.0, 0.0, 0.0) for _ in range(3)]
    # the walks read layers 0, 1 (forward) or 2, 1 (backward); the layer
    # they never read leads elsewhere
    steps = [_const_field(w, h, step_ts, dy, dx, 0.0) for _ in range(3)]
    steps[2 if forward else 0] = _const_field(w, h, step_ts, 5.0, -4.0, 0.0)
    cumul = (zero, steps) if forward else (steps, zero)
    rig = Rig(o, [{"color": P}], layers, r=r, cumul=cumul)
    return rig, dx, dy, 0 if forward else 2, far


@pytest.mark.parametrize("forward", [True, False])
def test_propagation_walks_follow_the_cumulative_field(forward):
    """From offsets at each pixel's own position in the near key, every
    neighbour candidate is followed through the field: the far key is
    reached at (x + 2dx, y + 2dy), where it matches the image exactly
    (residual 0), so interior pixels end there, in the far key's layer, with
    scale (1, 1).  Swapping the field's x and y channels, reading another
    field layer or walking the other way does not reach the match."""
    with OracleContext() as o:
        rig, dx, dy, near, far = _walk_rig(o, forward)
        rig.load(rig.offsets(near))
        o.propagate(0, 1, -1)
        off, _, res = rig.out()
    r = 1
    n = 0
    for y in range(r, rig.h - r):
        for x in range(r, rig.w - r):
            ex, ey = x + 2 * dx, y + 2 * dy
            if not (r <= ex < rig.sw - 1 - r and r <= ey < rig.sh - 1 - r):
                continue
            n += 1
            p = off[y, x]
            assert p["layer"] == far and p["x"] == ex and p["y"] == ey, (x, y, p)
            assert p["scale_u"] == 1.0 and p["scale_v"] == 1.0 and p["hysteresis"] == 0.0
            assert res[y, x] == 0.0
    assert n > 200


@pytest.mark.parametrize("ts,stops", [(-1.0, True), (-2e-6, True), (-5e-7, False), (0.0, False)])
def test_propagation_walk_stops_below_minus_epsilon(ts, stops):
    """propagate_kernel_keys stops a walk on a field texel whose time-step
    channel is below -EPSILON (1e-6): no pixel then leaves the near key."""
    with OracleContext() as o:
        rig, dx, dy, near, far = _walk_rig(o, True, step_ts=ts)
        rig.load(rig.offsets(near))
        o.propagate(0, 1, -1)
        off, _, _ = rig.out()
    n_far = int((off["layer"] == far).sum())
    if stops:
        assert set(np.unique(off["layer"])) == {near}
    else:
        assert n_far > 200


def test_scatter_walk_stops_on_an_invalid_ribbon_not_a_negative_step():
    """scatter_kernel_keys reads the field as RibbonP and stops on layer < 0
    only: with every time step negative its walks still run (pixels reach
    the other keys) unless the texels' layer channel is -1; propagation
    under the same negative steps never leaves the near key.  (The invalid
    texels carry layer -2: bilinear weights in FP32 may sum to just below
    1, and -1 * 0.99999994 would truncate to the valid layer 0.)"""
    counts = {}
    for lay in (1.0, -2.0):
        with OracleContext() as o:
            w, h = 40, 32
            layers = [{"base": W.sin_noise(21, k, w, h), "output": W.sin_noise(21, k, w, h)} for k in range(3)]
            fld = [_const_field(w, h, -1.0, 1.5, 2.0, lay) for _ in range(3)]
            rig = Rig(o, [{"color": W.sin_noise(22, 0, w, h)}], layers, cumul=(fld, fld),
                      rng_mode=_ffi.PT_RNG_COUNTER, scatter_samples_count=12)
            rig.load(rig.offsets(0))
            o.scatter(0, -1)
            off, _, _ = rig.out()
            counts[lay] = int((off["layer"] != 0).sum())
            rig.load(rig.offsets(0))
            o.propagate(0, 1, -1)
            off, _, _ = rig.out()
            assert set(np.unique(off["layer"])) == {0}
    assert counts[1.0] > 20 and counts[-2.0] == 0


def test_keyed_scatter_shares_samples_and_has_no_averaged_candidate():
    """The samples are divided among the keys (integer division, :765) and
    the averaged-neighbour candidate is absent: with 2 samples and 3 keys a
    keyed scatter tests nothing and changes nothing, while the plain scatter
    of the same state moves pixels; the PDE count is lw * lh * n_s for the
    shared sample count and no averaged candidate (plus lw * lh for the
    residual refresh of the sync that follows)."""
    w, h = 32, 24
    layers = [{"base": W.sin_noise(31, k, w, h), "output": W.sin_noise(31, k, w, h)} for k in range(3)]
    fr = [{"color": W.sin_noise(32, 0, w, h)}]
    with OracleContext() as o:
        rig = Rig(o, fr, layers, scatter_samples_count=2)
        rig.load(rig.offsets(0, 0.5, 0.25))
        rig.settle()
        before = rig.out()[0].copy()
        p0 = o.pde_count()
        o.scatter(0, -1)
        assert o.pde_count() - p0 == w * h
        assert np.array_equal(rig.out()[0], before)
    with OracleContext() as o:
        rig = Rig(o, fr, layers, keys=False, scatter_samples_count=2)
        rig.load(rig.offsets(0, 0.5, 0.25))
        rig.settle()
        p0 = o.pde_count()
        o.scatter(0, -1)
        # n_s(2) at exemplar 32 is 2, plus the averaged candidate
        assert o.pde_count() - p0 == w * h * (3 + 1)
        assert not np.array_equal(rig.out()[0], before)
    with OracleContext() as o:
        rig = Rig(o, fr, layers, scatter_samples_count=9)
        rig.load(rig.offsets(0, 0.5, 0.25))
        p0 = o.pde_count()
        o.scatter(0, -1)
        # 9 // 3 = 3 samples; 32 * 0.9^3 >= 10, so all three run
        assert o.pde_count() - p0 == w * h * (3 + 1)


def test_keymask_fixes_painted_pixels():
    """A keyframe index k: where style layer k's output is painted at the
    pixel's own position (alpha > 0, the texel itself), propagation sets
    PatchXF(x, y, 0, 1, k) and scatter keeps the current offset; elsewhere
    the operators search as usual."""
    w, h = 32, 28
    layers = []
    for k in range(3):
        out = W.sin_noise(41, k, w, h)
        out[..., 3] = np.where(W.smooth_noise(42, k, w, h) > 0.5, 0.75, 0.0)
        layers.append({"base": W.sin_noise(43, k, w, h), "output": out})
    painted = layers[1]["output"][..., 3] > 0
    with OracleContext() as o:
        rig = Rig(o, [{"color": W.sin_noise(44, 0, w, h)}], layers, rng_mode=_ffi.PT_RNG_COUNTER)
        start = rig.offsets(2, 0.5, 0.5)
        rig.load(start)
        o.propagate(0, 2, 1)
        off, _, _ = rig.out()
        yy, xx = np.mgrid[0:h, 0:w]
        assert painted.sum() > 100 and (~painted).sum() > 100
        k = off[painted]
        assert np.all(k["x"] == xx[painted]) and np.all(k["y"] == yy[painted]) and np.all(k["layer"] == 1)
        assert np.all(k["hysteresis"] == 1.0) and np.all(k["theta"] == 0.0) and np.all(k["scale_u"] == 1.0)
        assert not np.any((off[~painted]["hysteresis"] == 1.0) & (off[~painted]["layer"] == 1))
        rig.load(start)
        o.scatter(0, 1)
        off2, _, _ = rig.out()
        assert np.array_equal(off2[painted], start[painted])
        assert np.any(off2[~painted] != start[~painted])


def test_residual_and_canvas_sample_the_offsets_layer():
    """An offset samples the exemplar layer it carries, clamped into the
    layers as a layered texture does: the layer matching the image gives a
    zero residual, the others do not; the canvas is that layer's output."""
    w, h = 24, 20
    P = W.sin_noise(51, 0, w, h)
    layers = [{"base": W.sin_noise(52, k, w, h), "output": W.sin_noise(53, k, w, h)} for k in range(3)]
    layers[2]["base"] = P
    with OracleContext() as o:
        rig = Rig(o, [{"color": P}], layers, keys=False)
        for lay, zero in ((0, False), (1, False), (2, True), (7, True)):
            rig.load(rig.offsets(lay))
            off, canvas, res = rig.out()
            assert np.all(off["layer"] == lay)
            inner = res[1:-1, 1:-1]
            assert np.all(inner == 0.0) if zero else np.all(inner > 0.0)
            assert np.array_equal(canvas, layers[min(lay, 2)]["output"])


def test_random_style_draw_stays_within_the_keys():
    """With keys first..last each pixel's style is drawn among them (the
    texel's w, which randomisation turns into the offset's layer) and its
    position inside that style's exemplar, away from its last row and
    column (rand() % (size - 1))."""
    w, h = 40, 36
    with OracleContext() as o:
        prm = o.default_params(residual_window_size=1, levels=1, num_styles=3, interpolate_key_frame=1,
                               first_key_index=1, last_key_index=2, orientation_enabled=0)
        o.setup(prm)
        o.initialize(w, h, 30, 20)
        o.initialize_exemplars(30, 20, 3)
        col = o.pyramid(E4, W.sin_noise(61, 0, w, h))
        ori = o.pyramid(E1, None, shape=(w, h))
        sizes = [(30, 20), (16, 12), (8, 30)]
        outs = [o.pyramid(E4, np.ones((sh, sw, 4), np.float32)) for sw, sh in sizes]
        oris = [o.pyramid(E1, None, shape=(sw, sh)) for sw, sh in sizes]
        o.create_random_pyramid_keys(0, 3, col, ori, outs, oris)
        o.upload_image(0, "INPUT_ORIENTATION", ori)
        o.upload_image(0, "INPUT_SCALE", o.pyramid(E2, None, shape=(w, h), default=1.0))
        for k in range(3):
            o.upload_image(0, "EXEMPLAR_BASE", outs[0], k)
        o.randomize_offsets(0)
        off, _, _ = o.get_output(0, w, h)
    lay = off["layer"]
    assert set(np.unique(lay)) == {1, 2}
    assert 0.35 < (lay == 1).mean() < 0.65
    for k in (1, 2):
        sw, sh = sizes[k]
        sel = off[lay == k]
        assert sel["x"].min() >= 0 and sel["x"].max() < sw - 1 and sel["y"].max() < sh - 1
    assert off[lay == 2]["y"].max() > 12  # style 2's exemplar, not style 1's


def _constant_motion(w, h, vx, vy):
    """A frame pair moving by (vx, vy): the second frame's backward velocity
    and the first frame's forward velocity are consistent."""
    f0 = {"color": W.sin_noise(71, 0, w, h), "vel_f": np.tile(np.float32([vx, vy]), (h, w, 1)),
          "vel_b": np.zeros((h, w, 2), np.float32)}
    f1 = {"color": W.sin_noise(71, 1, w, h), "vel_f": np.zeros((h, w, 2), np.float32),
          "vel_b": np.tile(np.float32([-vx, -vy]), (h, w, 1))}
    return [f0, f1]


def _ribbons(w, h, x, y, ts, layer):
    r = np.zeros((h, w), _ffi.RIBBON_DTYPE)
    r["x"], r["y"], r["time_step"], r["layer"] = x, y, ts, layer
    return r


def _advance(rig):
    rig.o.flip_anim_buffers()
    rig.load_frame(1, random=False)


def test_ribbon_accumulation_integer_motion():
    """With a constant integer motion (2, 1) every ribbon moves with its
    pixel: out(x, y) = in(x - 2, y - 1) + (-2, -1) and the time step grows
    by time_step; pixels whose source lies outside the frame become invalid
    when overwriting and are kept otherwise, and a kept ribbon is replaced
    only by one whose time step is not longer."""
    w, h = 20, 16
    layers = [{"base": W.sin_noise(72, k, w, h), "output": W.sin_noise(72, k, w, h)} for k in range(2)]
    with OracleContext() as o:
        rig = Rig(o, _constant_motion(w, h, 2.0, 1.0), layers)
        o.randomize_offsets(0)
        _advance(rig)
        src = _ribbons(w, h, np.arange(w)[None, :] * 0.25, 0.5, 3, 1)
        o.load_ribbon_field(0, src)
        o.accumulate_ribbon_field(0, True, 2, True)
        got = o.get_ribbon_field(0, w, h)
        for y in range(h):
            for x in range(w):
                g = got[y, x]
                if x >= 2 and y >= 1:
                    assert g["layer"] == 1 and g["time_step"] == 5, (x, y, g)
                    assert g["x"] == np.float32(src[y - 1, x - 2]["x"]) - 2 and g["y"] == np.float32(0.5) - 1
                else:
                    assert g["layer"] == -1
        # not overwriting: out-of-frame sources keep the old ribbon; a longer
        # ribbon (5 + 1 = 6 > 5) never replaces a valid one
        o.accumulate_ribbon_field(0, True, 1, False)
        kept = o.get_ribbon_field(0, w, h)
        assert np.array_equal(kept, got)
        # one of equal length does: sources of time step 4 arrive with |4 + 1| = 5 <= 5
        o.load_ribbon_field(0, _ribbons(w, h, 0.0, 0.0, 4, 0))
        o.accumulate_ribbon_field(0, True, 1, False)
        again = o.get_ribbon_field(0, w, h)
        assert np.all(again[1:, 2:]["time_step"] == 5) and np.all(again[1:, 2:]["layer"] == 0)
        assert np.all(again[1:, 2:]["x"] == -2) and np.all(again[1:, 2:]["y"] == -1)
        assert np.array_equal(again[:1], got[:1]) and np.array_equal(again[:, :2], got[:, :2])


def test_ribbon_accumulation_interpolates_only_when_the_chained_layer_test_holds():
    """At a fractional source position the four ribbons around it are
    interpolated bilinearly when all are valid, close (5 px) and pass the
    reference's chained test (l00 == l10 == l01 == l11, evaluated left to
    right: true for layers 0 and 1, false for layer 2); otherwise the
    nearest one is taken."""
    w, h = 20, 16
    layers = [{"base": W.sin_noise(73, k, w, h), "output": W.sin_noise(73, k, w, h)} for k in range(3)]
    xs = (np.arange(w)[None, :] * 0.5 + np.arange(h)[:, None] * 0.25).astype(np.float32)
    for lay, interp in ((0, True), (1, True), (2, False)):
        with OracleContext() as o:
            rig = Rig(o, _constant_motion(w, h, 1.5, 0.5), layers)
            _advance(rig)
            o.load_ribbon_field(0, _ribbons(w, h, xs, 0.0, 1, lay))
            o.accumulate_ribbon_field(0, True, 1, True)
            got = o.get_ribbon_field(0, w, h)
        for y in range(2, h - 1):
            for x in range(3, w - 1):
                tx, ty = x - 1.5, y - 0.5
                fx, fy = int(np.floor(tx)), int(np.floor(ty))
                a, b = tx - fx, ty - fy
                q = xs.astype(np.float64)
                if interp:
                    e = ((1 - a) * q[fy, fx] + a * q[fy, fx + 1]) * (1 - b) + ((1 - a) * q[fy + 1, fx] +
                                                                             a * q[fy + 1, fx + 1]) * b
                else:
                    # the nearest ribbon: round half away from zero (roundf)
                    e = q[int(np.floor(ty + 0.5)), int(np.floor(tx + 0.5))]
                assert abs(got[y, x]["x"] - (e - 1.5)) < 1e-5, (lay, x, y)
                assert got[y, x]["layer"] == lay and got[y, x]["time_step"] == 2


def test_ribbon_canvas_lookup():
    """After accumulation the forward advected output colour of a pixel with
    a valid ribbon is the canvas texel at (int(x + rx), int(y + ry)),
    clamped into the canvas, and zero where the ribbon is invalid."""
    w, h = 20, 16
    layers = [{"base": W.sin_noise(74, k, w, h), "output": W.sin_noise(75, k, w, h)} for k in range(2)]
    rng = np.random.default_rng(9)
    with OracleContext() as o:
        rig = Rig(o, _constant_motion(w, h, 0.0, 0.0), layers)
        rig.load(rig.offsets(0, 0.25, 0.5))
        _advance(rig)
        src = _ribbons(w, h, rng.uniform(-6, 6, (h, w)), rng.uniform(-6, 6, (h, w)), 1,
                       np.where(rng.uniform(0, 1, (h, w)) < 0.3, -1, 0))
        o.load_ribbon_field(0, src)
        o.accumulate_ribbon_field(0, True, 0, True)
        rib = o.get_ribbon_field(0, w, h)
        _, canvas, _ = o.get_output(0, w, h)
        _, _, out, _, _ = o.get_advected(0, 0, w, h)
    v = src["layer"] >= 0  # zero motion: every valid ribbon stays, the others become RibbonP()
    assert np.array_equal(rib[v], src[v]) and np.all(rib[~v]["layer"] == -1)
    for y in range(h):
        for x in range(w):
            if rib[y, x]["layer"] < 0:
                assert np.all(out[y, x] == 0)
            else:
                sx = min(max(int(np.float32(x) + rib[y, x]["x"]), 0), w - 1)
                sy = min(max(int(np.float32(y) + rib[y, x]["y"]), 0), h - 1)
                assert np.array_equal(out[y, x], canvas[sy, sx]), (x, y)


def _ftk(w, h, seed, nk):
    rng = np.random.default_rng(seed)
    t = np.zeros((h, w, 4), np.float32)
    t[..., 0] = rng.integers(-4, 5, (h, w))
    t[..., 1] = rng.integers(-3, 4, (h, w))
    t[..., 2] = rng.integers(-3, 4, (h, w))
    t[..., 3] = np.where(rng.uniform(0, 1, (h, w)) < 0.3, -1, rng.integers(0, nk, (h, w)))
    return t


def test_offsets_from_frame_to_key_ribbons():
    """Each pixel starts at the keyframe texel its ribbon reaches,
    (x + rx, y + ry) in the ribbon's layer with hysteresis 1, from the F
    ribbons when looking forwards and the B ribbons otherwise; an invalid
    ribbon gives the random texel (hysteresis 0)."""
    w, h = 24, 20
    layers = [{"base": W.sin_noise(81, k, w, h), "output": W.sin_noise(81, k, w, h)} for k in range(3)]
    fb, ff = _ftk(w, h, 1, 3), _ftk(w, h, 2, 3)
    with OracleContext() as o:
        rig = Rig(o, [{"color": W.sin_noise(82, 0, w, h), "FRAME_TO_KEY_RIBBON_B": fb,
                       "FRAME_TO_KEY_RIBBON_F": ff}], layers)
        for look, t in ((True, ff), (False, fb)):
            o.offsets_from_frame_to_key_ribbon(0, look)
            off, _, _ = rig.out()
            yy, xx = np.mgrid[0:h, 0:w]
            v = t[..., 3] >= 0
            assert np.array_equal(off["x"][v], xx[v] + t[..., 2][v])
            assert np.array_equal(off["y"][v], yy[v] + t[..., 1][v])
            assert np.array_equal(off["layer"][v], t[..., 3][v].astype(np.int32))
            assert np.all(off["hysteresis"][v] == 1.0) and np.all(off["hysteresis"][~v] == 0.0)
            assert np.all(off["layer"][~v] == 0) and off["x"][~v].max() < w


def test_advection_frame_to_key_check():
    """ADVECT_CHECK_FRAME_TO_KEY with a long step (3): where the frame's B
    ribbon (advecting forwards) is valid and shorter than the step, the
    advected offset becomes (x + rx, y + ry) in the ribbon's layer; in every
    checked pixel the advected colours become that offset's exemplar texels
    (base and output of its layer) and the time step |ribbon time step|.
    With a step of 1 and a valid advection the check does not run."""
    w, h = 24, 20
    layers = [{"base": W.sin_noise(91, k, w, h), "output": W.sin_noise(92, k, w, h)} for k in range(3)]
    fb = _ftk(w, h, 3, 3)
    frames = _constant_motion(w, h, 1.0, 0.0)
    frames[1]["FRAME_TO_KEY_RIBBON_B"] = fb
    with OracleContext() as o:
        rig = Rig(o, frames, layers)
        rig.load(rig.offsets(0, 0.0, 0.0))
        _advance(rig)
        o.advect_offsets(0, 3, True, CHECK)
        adv, base, out, _, ts = o.get_advected(0, 0, w, h)
        n_key = n_adv = 0
        for y in range(h):
            for x in range(w):
                rb = fb[y, x]
                valid_adv = x >= 1
                a = adv[y, x]
                if rb[3] >= 0 and (abs(rb[0]) < 3 or not valid_adv):
                    n_key += 1
                    assert (a["x"], a["y"], a["layer"]) == (x + rb[2], y + rb[1], rb[3]), (x, y)
                elif valid_adv:
                    n_adv += 1
                    # the advected offset: the identity offset of the source (x - 1, y)
                    assert (a["x"], a["y"], a["layer"]) == (x - 1, y, 0), (x, y)
                else:
                    assert a["hysteresis"] == 0  # random
                    continue
                assert ts[y, x] == abs(int(rb[0]))
                ex = min(max(int(a["x"]), 0), w - 1), min(max(int(a["y"]), 0), h - 1)
                lay = layers[int(a["layer"])]
                assert np.array_equal(base[y, x], lay["base"][ex[1], ex[0]])
                assert np.array_equal(out[y, x], lay["output"][ex[1], ex[0]])
        assert n_key > 50 and n_adv > 50
        # a single step with a valid advection: no check, the previous frame's colour
        rig.load(rig.offsets(0, 0.0, 0.0))
        o.advect_offsets(0, 1, True, CHECK)
        adv, base, _, _, ts = o.get_advected(0, 0, w, h)
        c0 = frames[0]["color"]
        assert np.all(ts[:, 1:] == 1) and np.array_equal(base[:, 1:], c0[:, :-1])


def test_merge_randomly_prefers_the_advected_offset_where_only_its_ribbon_is_valid():
    """mergeRandomly with weight 0: the advected offset wins exactly where
    this direction's frame-to-key ribbon (B when merging forwards) is valid
    and the other one is not."""
    w, h = 24, 20
    layers = [{"base": W.sin_noise(93, k, w, h), "output": W.sin_noise(93, k, w, h)} for k in range(2)]
    fb, ff = _ftk(w, h, 4, 2), _ftk(w, h, 5, 2)
    frames = _constant_motion(w, h, 1.0, 0.0)
    frames[1].update(FRAME_TO_KEY_RIBBON_B=fb, FRAME_TO_KEY_RIBBON_F=ff)
    with OracleContext() as o:
        rig = Rig(o, frames, layers)
        rig.load(rig.offsets(1, 0.5, 0.5))
        _advance(rig)
        o.advect_offsets(0, 1, True, 0)
        adv = o.get_advected(0, 0, w, h)[0]
        rig.load(rig.offsets(0, 2.0, 1.0))
        base = rig.out()[0].copy()
        o.merge_randomly(0, True, 0.0)
        got = rig.out()[0]
    pick = (fb[..., 3] >= 0) & ~(ff[..., 3] >= 0)
    assert pick.sum() > 20 and (~pick).sum() > 20
    assert np.array_equal(got[pick], adv[pick]) and np.array_equal(got[~pick], base[~pick])


def test_merge_advected_buffers_cases():
    """Both advected offsets valid: one of them, each chosen about half the
    time; only the forward one valid: it; neither: the fractional random
    texel as an offset (x, y in [0, 1), layer 0, hysteresis 0) with
    MERGE_FALLBACK_RANDOM, the current offset with MERGE_FALLBACK_LAST_PASS."""
    w, h = 24, 20
    layers = [{"base": W.sin_noise(94, k, w, h), "output": W.sin_noise(94, k, w, h)} for k in range(2)]
    frames = _constant_motion(w, h, 1.0, 0.0)
    # backward advection from the second frame: sources at (x + 1, y)
    frames[1]["vel_f"] = np.tile(np.float32([1.0, 0.0]), (h, w, 1))
    frames[0]["vel_b"] = np.tile(np.float32([-1.0, 0.0]), (h, w, 1))
    nocheck = _ffi.PT_ADVECT_DO_NOT_CHECK_PREVIOUS
    with OracleContext() as o:
        rig = Rig(o, frames, layers)
        rig.load(rig.offsets(1, 0.5, 0.5))
        _advance(rig)
        o.advect_offsets(0, 1, True, 0)
        o.advect_offsets(0, 1, False, nocheck)  # every backward offset invalid
        af, ab = o.get_advected(0, 0, w, h)[0], o.get_advected(0, 1, w, h)[0]
        o.merge_advected_buffers(0, _ffi.PT_MERGE_FALLBACK_RANDOM)
        got = rig.out()[0]
        vf = af["hysteresis"] > 0
        assert vf.sum() > 100 and np.all(ab["hysteresis"] == 0)
        assert np.array_equal(got[vf], af[vf])
        nv = ~vf
        assert np.all(got[nv]["x"] >= 0) and np.all(got[nv]["x"] < 1) and np.all(got[nv]["y"] < 1)
        assert np.all(got[nv]["layer"] == 0) and np.all(got[nv]["hysteresis"] == 0)
        # neither valid, last pass: the current offsets stay
        rig.load(rig.offsets(1, 3.0, 2.0))
        cur = rig.out()[0].copy()
        o.advect_offsets(0, 1, True, nocheck)
        rig.load(cur)
        o.merge_advected_buffers(0, _ffi.PT_MERGE_FALLBACK_LAST_PASS)
        assert np.array_equal(rig.out()[0], cur)
        # both valid: each pixel takes one of the two, both occur
        rig.load(rig.offsets(1, 0.5, 0.5))
        o.advect_offsets(0, 1, True, 0)
        o.advect_offsets(0, 1, False, 0)
        af, ab = o.get_advected(0, 0, w, h)[0], o.get_advected(0, 1, w, h)[0]
        o.merge_advected_buffers(0, _ffi.PT_MERGE_FALLBACK_LAST_PASS)
        got = rig.out()[0]
    both = (af["hysteresis"] > 0) & (ab["hysteresis"] > 0) & (af["x"] != ab["x"])
    assert both.sum() > 100
    is_f = got[both] == af[both]
    is_b = got[both] == ab[both]
    assert np.all(is_f | is_b) and 0.3 < is_f.mean() < 0.7


def test_ribbon_pyramid_closed_form():
    """Level L+1 of a ribbon pyramid is the top-left texel of each 2x2 block
    of level L as a RibbonP: time step and layer truncated toward zero, x
    and y halved."""
    rng = np.random.default_rng(12)
    t = np.zeros((16, 20, 4), np.float32)
    t[..., 0] = rng.uniform(-5, 5, t.shape[:2])
    t[..., 1] = rng.uniform(-8, 8, t.shape[:2])
    t[..., 2] = rng.uniform(-8, 8, t.shape[:2])
    t[..., 3] = rng.uniform(-1.5, 3, t.shape[:2])
    with OracleContext() as o:
        p = o.pyramid(E4, t, RIBBON)
        lv = [p.download(l) for l in range(p.levels)]
    assert np.array_equal(lv[0], t)
    for l in range(1, len(lv)):
        prev = lv[l - 1][::2, ::2][:lv[l].shape[0], :lv[l].shape[1]]
        exp = np.stack([np.trunc(prev[..., 0]), prev[..., 1] * np.float32(0.5), prev[..., 2] * np.float32(0.5),
                        np.trunc(prev[..., 3])], -1).astype(np.float32)
        assert np.array_equal(lv[l], exp), l
