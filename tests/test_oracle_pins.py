"""Pins of the CPU oracle against what the mathematics fixes, so a plausible
mistake in the restatement (a dropped term, a wrong sign or index, a
transposed operand, a missing wrap) fails here.  The reference ships no tests,
fixtures or golden data (see the oracle's header), so these checks, not the
reference's outputs, anchor the oracle."""
import numpy as np
import pytest

from painttween_amd import _ffi
from painttween_amd import workloads as W
from tests.harness import Session
from tests.oracle_ctx import OracleContext, load_oracle

E1, E2, E3, E4 = 1, 2, 3, 4


# ---------------------------------------------------------------- weights
def test_gaussian_weights_printed_values():
    """allocateGaussianKernel(r, false) (convolution.cu:37-60):
    k[i] = exp(-i^2 / (2 sigma^2)), sigma = r/3, unnormalised."""
    with OracleContext() as o:
        o.setup(o.default_params(residual_window_size=3))
        k = o.gaussian_weights(3)
        np.testing.assert_allclose(k, [0.0111090, 0.135335, 0.606531, 1, 0.606531, 0.135335, 0.0111090], rtol=1e-5)
        np.testing.assert_allclose(k, np.exp(-np.arange(-3, 4) ** 2 / 2.0), rtol=1e-6)
        o.setup(o.default_params(residual_window_size=2))
        k = o.gaussian_weights(2)
        np.testing.assert_allclose(k, [0.0111090, 0.324652, 1, 0.324652, 0.0111090], rtol=1e-5)
        np.testing.assert_allclose(k, np.exp(-np.arange(-2, 3) ** 2 / (2 * (2 / 3) ** 2)), rtol=1e-6)


# ---------------------------------------------------------------- pyramid
def _box(prev):
    h, w = prev.shape[0] // 2, prev.shape[1] // 2
    p = prev[:2 * h, :2 * w].astype(np.float32)
    s = np.float32(0) + p[0::2, 0::2]
    s = s + p[0::2, 1::2]
    s = s + p[1::2, 0::2]
    s = s + p[1::2, 1::2]
    return (s * np.float32(0.25)).astype(np.float32)


def test_box_pyramid_closed_form_and_odd_sizes():
    a = np.random.default_rng(1).uniform(0, 1, (37, 53, 4)).astype(np.float32)
    with OracleContext() as o:
        p = o.pyramid(E4, a)
        assert p.levels == 5  # min(floor(log2 53), floor(log2 37))
        prev = a
        for l in range(1, p.levels):
            exp = _box(prev)
            assert p.dims(l) == (53 >> l, 37 >> l)
            got = p.download(l)
            assert got.tobytes() == exp.tobytes(), l
            prev = exp


def test_pyramid_constant_and_velocity_scaling():
    with OracleContext() as o:
        c = o.pyramid(E1, np.full((32, 32), 0.375, np.float32))
        for l in range(c.levels):
            assert np.all(c.download(l) == np.float32(0.375))
        v = o.pyramid(E2, np.full((32, 32, 2), 3.0, np.float32), _ffi.PT_DOWNSAMPLE_SCALE_WITH_LEVELS)
        for l in range(v.levels):
            assert np.all(v.download(l) == np.float32(3.0 * 0.5 ** l)), l


def test_absent_inputs_use_defaults():
    with OracleContext() as o:
        d = o.pyramid(E1, None, default=1.0, shape=(16, 8))
        assert d.dims(0) == (16, 8) and np.all(d.download(0) == 1.0)
        t = o.pyramid(E3, None, shape=(16, 8))
        a = o.orientation_pyramid(t)
        assert np.all(a.download(1) == np.float32(float(np.float32(np.pi / 2)) - np.pi))


# ---------------------------------------------------------------- patch distance
def _bilinear(tex, x, y):
    h, w = tex.shape[:2]
    fx, fy = np.floor(x), np.floor(y)
    a, b = x - fx, y - fy
    x0, y0 = int(fx), int(fy)
    xa, xb = np.clip([x0, x0 + 1], 0, w - 1)
    ya, yb = np.clip([y0, y0 + 1], 0, h - 1)
    t = tex.astype(np.float64)
    return ((1 - a) * (1 - b) * t[ya, xa] + a * (1 - b) * t[ya, xb] + (1 - a) * b * t[yb, xa] + a * b * t[yb, xb])


def _coh_double(cx, cy, tcx, tcy, tx, ty, i, j, level, rr, ar):
    """coherenceTerm (texSynth_kernel.cu:377-423) in float64."""
    dx, dy = cx - tcx, cy - tcy
    ratio = min(max((dx * dx + dy * dy) / float(i * i + j * j), 1.0 / rr), rr)
    er = np.log(ratio) ** 2 / np.log(rr) ** 2
    ux, uy = tx - tcx, ty - tcy
    lt = np.hypot(ux, uy)
    ux, uy = (ux / lt, uy / lt) if lt > 1e-4 else (1.0, 0.0)
    lc = np.hypot(dx, dy)
    vx, vy = (dx / lc, dy / lc) if lc > 1e-4 else (1.0, 0.0)
    ang = min(np.arccos(np.clip(ux * vx + uy * vy, -1.0, 1.0)), 1.0)
    ea = min(ang, ar * np.pi) / (ar * np.pi)
    return 1.0 - (1.0 - er) * (1.0 - ea)


def _pde_double(x, y, cand, target, canvas, exb, exo, exdt, imgdt, g, prm, off=None, level=0, ids=None, styles=1):
    """Independent float64 statement of calculateResidual
    (texSynth_kernel.cu:537-659) without temporal terms; the coherence term
    (:642-646) against the offset field `off` when its weight is positive; the
    object-ID term (:497-504) with ids = (image IDs, exemplar IDs) when
    style_obj_id_weight is positive."""
    r = prm.residual_window_size
    lh, lw = target.shape[:2]
    hh, hw = exb.shape[:2]
    hh *= styles  # the wrap height is the offsets histogram's: style height x numStyles
    th = float(cand["theta"])
    su = prm.residual_window_scale_u * float(cand["scale_u"])
    sv = prm.residual_window_scale_v * float(cand["scale_v"])
    b0 = (np.cos(th) * su, np.sin(th) * su)
    b1 = (-np.sin(th) * sv, np.cos(th) * sv)
    num = den = 0.0
    for i in range(-r, r + 1):
        for j in range(-r, r + 1):
            if not (0 <= x + i < lw and 0 <= y + j < lh):
                continue
            tx = float(cand["x"]) + i * b0[0] + j * b1[0]
            ty = float(cand["y"]) + i * b0[1] + j * b1[1]
            if tx >= hw:
                tx -= hw
            elif tx < 0:
                tx += hw
            if ty >= hh:
                ty -= hh
            elif ty < 0:
                ty += hh
            tb, to = _bilinear(exb, tx, ty), _bilinear(exo, tx, ty)
            d = target[y + j, x + i].astype(np.float64) - tb
            res = (d[0] ** 2 + d[1] ** 2 + d[2] ** 2 + prm.alpha_weight * d[3] ** 2) * prm.input_analogy_weight
            if i == 0 and j == 0:
                if prm.dist_trans_weight > 0:
                    de = _bilinear(exdt[..., None], tx, ty)[0] / float(cand["scale_v"])
                    res += min(max(abs(de - float(imgdt[y, x])), 0.0), 10.0) * prm.dist_trans_weight
                if prm.style_obj_id_weight > 0:
                    # point sampling at coordinate + 0.5, clamped (tex2D, cudaFilterModePoint)
                    ex = min(max(int(np.floor(tx + 0.5)), 0), hw - 1)
                    ey = min(max(int(np.floor(ty + 0.5)), 0), hh - 1)
                    if ids[0][y, x] != ids[1][ey, ex]:
                        res += prm.style_obj_id_weight
            else:
                d = canvas[y + j, x + i].astype(np.float64) - to
                res += (d[0] ** 2 + d[1] ** 2 + d[2] ** 2 + prm.alpha_weight * d[3] ** 2) * prm.canvas_output_weight
                if prm.coherence_weight > 0:
                    n = off[y + j, x + i]
                    res += _coh_double(float(n["x"]), float(n["y"]), float(cand["x"]), float(cand["y"]), tx, ty, i,
                                       j, level, prm.coherence_ratio_range,
                                       prm.coherence_angle_range) * prm.coherence_weight
            wgt = float(g[i + r]) * float(g[j + r])
            num += res * wgt
            den += wgt
    return num / den


@pytest.mark.parametrize("r,level", [(2, 0), (3, 0), (3, 1)])
def test_patch_distance_matches_float64_brute_force(r, level):
    st = W.style_pair(24, 20, seed=41, guides=True)
    fr = W.still_frame(22, 18, seed=42, guides=True)
    rng = np.random.default_rng(43)
    with OracleContext() as o:
        prm = o.default_params(residual_window_size=r, levels=2, dist_trans_weight=0.7, alpha_weight=1.5,
                               canvas_output_weight=0.8, input_analogy_weight=1.3, residual_window_scale_u=1.1,
                               residual_window_scale_v=0.9)
        s = Session(o, st, [fr], prm)
        s.run([("style", level), ("frame", level, 0), ("randomize",)])
        off, canvas, _ = s.state()
        exb, exo = s.sb.download(level), s.so.download(level)
        exdt = s.sdt.download(level)
        target = s.frames[0]["color"].download(level)
        imgdt = s.frames[0]["dt"].download(level)
        g = o.gaussian_weights(r)
        lh, lw = target.shape[:2]
        pts = [(0, 0), (lw - 1, lh - 1), (0, lh - 1), (lw - 1, 0)] + \
              [(int(a), int(b)) for a, b in zip(rng.integers(0, lw, 12), rng.integers(0, lh, 12))]
        for x, y in pts:
            cand = np.zeros((), _ffi.PATCHXF_DTYPE)
            cand["x"] = rng.uniform(-3, exb.shape[1] + 3)
            cand["y"] = rng.uniform(-3, exb.shape[0] + 3)
            cand["theta"] = rng.uniform(-1.5, 1.5)
            cand["scale_u"] = rng.uniform(0.8, 1.2)
            cand["scale_v"] = rng.uniform(0.8, 1.2)
            got = o.patch_distance(x, y, level, cand)
            exp = _pde_double(x, y, cand, target, canvas, exb, exo, exdt, imgdt, g, prm)
            assert abs(got - exp) <= 3e-6 * max(exp, 1e-3), (x, y, got, exp)


@pytest.mark.parametrize("r,level", [(2, 0), (3, 1)])
def test_patch_distance_with_coherence_matches_float64(r, level):
    """The coherence term against the float64 statement, random candidates
    over a random offset field (directions away from the acos edge)."""
    st = W.style_pair(24, 20, seed=44, guides=True)
    fr = W.still_frame(22, 18, seed=45, guides=True)
    rng = np.random.default_rng(46)
    with OracleContext() as o:
        prm = o.default_params(residual_window_size=r, levels=2, coherence_weight=0.6, coherence_ratio_range=2.5,
                               coherence_angle_range=0.3)
        s = Session(o, st, [fr], prm)
        s.run([("style", level), ("frame", level, 0), ("randomize",)])
        off, canvas, _ = s.state()
        exb, exo = s.sb.download(level), s.so.download(level)
        exdt = s.sdt.download(level)
        target = s.frames[0]["color"].download(level)
        imgdt = s.frames[0]["dt"].download(level)
        g = o.gaussian_weights(r)
        lh, lw = target.shape[:2]
        for x, y in [(0, 0), (lw - 1, lh - 1)] + [(int(a), int(b)) for a, b in
                                                  zip(rng.integers(0, lw, 10), rng.integers(0, lh, 10))]:
            cand = np.zeros((), _ffi.PATCHXF_DTYPE)
            cand["x"] = rng.uniform(2, exb.shape[1] - 2)
            cand["y"] = rng.uniform(2, exb.shape[0] - 2)
            cand["theta"] = rng.uniform(-1.5, 1.5)
            cand["scale_u"] = cand["scale_v"] = 1.0
            got = o.patch_distance(x, y, level, cand)
            exp = _pde_double(x, y, cand, target, canvas, exb, exo, exdt, imgdt, g, prm, off=off, level=level)
            assert abs(got - exp) <= 3e-5 * max(exp, 1e-3), (x, y, got, exp)


def test_coherence_term_special_values():
    """Consistent neighbour (same distance and direction) costs exactly 0; a
    neighbour at twice the distance (ratio clamped to the range) or in the
    opposite direction (angle clamped to 1 rad > range*pi) costs exactly 1."""
    lib = load_oracle()
    f = lib.ptref_coherence_term
    assert f(11.0, 10.0, 10.0, 10.0, 11.0, 10.0, 1, 0, 0, 2.0, 0.25) == 0.0
    assert f(10.0, 11.0, 10.0, 10.0, 10.0, 11.0, 0, 1, 1, 2.0, 0.25) == 0.0
    assert f(12.0, 10.0, 10.0, 10.0, 11.0, 10.0, 1, 0, 0, 2.0, 0.25) == 1.0
    assert f(9.0, 10.0, 10.0, 10.0, 11.0, 10.0, 1, 0, 0, 2.0, 0.25) == 1.0
    # a right angle with ratio 1: Eangle = min(pi/2 -> 1, 0.5*pi) / (0.5*pi)
    got = f(10.0, 11.0, 10.0, 10.0, 11.0, 10.0, 1, 0, 2, 2.0, 0.5)
    assert abs(got - 1.0 / (0.5 * np.pi)) < 1e-6, got


def _identity_nnf(w, h, dx=0, dy=0, ew=None, eh=None):
    ew, eh = ew or w, eh or h
    n = np.zeros((h, w), _ffi.PATCHXF_DTYPE)
    yy, xx = np.mgrid[0:h, 0:w]
    n["x"] = (xx + dx) % ew
    n["y"] = (yy + dy) % eh
    n["scale_u"] = 1
    n["scale_v"] = 1
    return n


def test_identity_offsets_give_zero_energy_and_copy_exemplar():
    """Target = exemplar input = exemplar output and the identity field: every
    tap compares a texel with itself, so the residual is exactly 0 and the
    canvas is the exemplar output exactly (offsetsToOutput, :263-274)."""
    ex = W.sin_noise(51, 0, 32, 24)
    with OracleContext() as o:
        s = Session(o, {"base": ex, "output": ex}, [{"color": ex}],
                    o.default_params(residual_window_size=3, levels=1, orientation_enabled=0))
        s.run([("style", 0), ("frame", 0, 0)])
        o.load_offsets(0, _identity_nnf(32, 24))
        o.sync_buffers(0, _ffi.PT_DONT_COPY_OFFSETS)
        _, canvas, res = s.state()
        assert canvas.tobytes() == ex.tobytes()
        assert np.all(res == 0.0)


def test_periodic_exemplar_integer_shift_wraps_to_zero():
    """A toroidally periodic exemplar, the target equal to its (dx, dy) shift,
    and the field x -> x + (dx, dy) mod size: the patch taps wrap by the
    exemplar size (texSynth_kernel.cu:590-599) and every difference vanishes."""
    w, h, dx, dy = 32, 32, 7, 29
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    ex = np.stack([0.5 + 0.4 * np.sin(2 * np.pi * (k + 1) * xx / w) * np.cos(2 * np.pi * (k + 2) * yy / h)
                   for k in range(4)], -1).astype(np.float32)
    tgt = np.roll(np.roll(ex, -dy, 0), -dx, 1)
    with OracleContext() as o:
        s = Session(o, {"base": ex, "output": ex}, [{"color": tgt}],
                    o.default_params(residual_window_size=3, levels=1, orientation_enabled=0))
        s.run([("style", 0), ("frame", 0, 0)])
        o.load_offsets(0, _identity_nnf(w, h, dx, dy))
        o.sync_buffers(0, _ffi.PT_DONT_COPY_OFFSETS)
        _, canvas, res = s.state()
        assert np.all(res == 0.0), res.max()


def test_bilinear_is_exact_on_affine_fields():
    yy, xx = np.mgrid[0:24, 0:32].astype(np.float32)
    ex = np.stack([xx * 0.25, yy * 0.5, xx * 0.125 + yy * 0.0625, np.full_like(xx, 1.0)], -1).astype(np.float32)
    nnf = _identity_nnf(16, 16)
    rng = np.random.default_rng(3)
    nnf["x"] = rng.uniform(0, 30.5, (16, 16)).astype(np.float32)
    nnf["y"] = rng.uniform(0, 22.5, (16, 16)).astype(np.float32)
    with OracleContext() as o:
        s = Session(o, {"base": ex, "output": ex}, [{"color": W.sin_noise(1, 2, 16, 16)}],
                    o.default_params(residual_window_size=1, levels=1, orientation_enabled=0))
        s.run([("style", 0), ("frame", 0, 0)])
        o.load_offsets(0, nnf)
        o.sync_buffers(0, _ffi.PT_DONT_COPY_OFFSETS | _ffi.PT_DONT_UPDATE_RESIDUAL)
        _, canvas, _ = s.state()
        np.testing.assert_allclose(canvas[..., 0], nnf["x"] * 0.25, atol=2e-6)
        np.testing.assert_allclose(canvas[..., 1], nnf["y"] * 0.5, atol=2e-6)


# ---------------------------------------------------------------- selection
def _seq(levels, iters):
    from tests.test_gpu_parity import _still_ops
    return _still_ops(levels, iters)


def test_selection_never_increases_energy_without_canvas_term():
    """With the canvas term off a pixel's residual depends on its own offset
    only, so the strict-min selection of propagate and scatter can never raise
    any pixel's residual."""
    st = W.style_pair(24, 24, seed=61)
    fr = W.still_frame(20, 20, seed=62)
    with OracleContext() as o:
        s = Session(o, st, [fr], o.default_params(residual_window_size=2, levels=1, canvas_output_weight=0.0))
        states = s.run(_seq(1, 3))
    prev = None
    for name, (_, _, res) in states:
        if prev is not None and name in ("propagate", "scatter"):
            assert np.all(res <= prev), name
        prev = res
    assert states[-1][1][2].mean() < 0.75 * states[4][1][2].mean()


def test_result_independent_of_thread_count():
    st = W.style_pair(40, 40, seed=71, guides=True)
    fr = W.still_frame(36, 30, seed=72, guides=True)
    outs = []
    for n in (1, 3, 8):
        with OracleContext(n) as o:
            outs.append(o.synthesize(o.default_params(residual_window_size=2, levels=2, dist_trans_weight=0.5),
                                     [fr], st, iterations=2))
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            assert a.tobytes() == b.tobytes()


# ---------------------------------------------------------------- upsample
def test_upsample_constant_field_and_border_ring():
    """A constant coarse field (cx, cy, theta 0) gives 2c + 0.5 inside the fine
    level; the one-pixel ring keeps the working buffer's content, which is
    still zero here (texSynth_kernel.cu:1461, cudaImageBuffer.h:332-334)."""
    with OracleContext() as o:
        s = Session(o, W.style_pair(32, 32, seed=81), [W.still_frame(32, 24, seed=82)],
                    o.default_params(residual_window_size=2, levels=2, orientation_enabled=0))
        s.run([("style", 0), ("frame", 0, 0)])
        coarse = _identity_nnf(16, 12)
        coarse["x"] = 5.25
        coarse["y"] = 7.5
        o.load_offsets(1, coarse)
        o.upsample(0)
        off, _, _ = o.get_output(0, 32, 24)
        inner = off[1:-1, 1:-1]
        assert np.all(inner["x"] == np.float32(2 * 5.25 + 0.5)) and np.all(inner["y"] == np.float32(2 * 7.5 + 0.5))
        ring = np.concatenate([off[0], off[-1], off[:, 0], off[:, -1]])
        assert np.all(ring["x"] == 0) and np.all(ring["y"] == 0) and np.all(ring["scale_u"] == 0)


# ---------------------------------------------------------------- advection
def _advect_setup(o, vel_f_prev, vel_b_cur, w=24, h=20):
    st = W.style_pair(32, 32, seed=91)
    f0 = {"color": W.sin_noise(92, 2, w, h), "vel_f": vel_f_prev, "vel_b": np.zeros((h, w, 2), np.float32)}
    f1 = {"color": W.sin_noise(93, 2, w, h), "vel_f": np.zeros((h, w, 2), np.float32), "vel_b": vel_b_cur}
    s = Session(o, st, [f0, f1], o.default_params(residual_window_size=2, levels=1, orientation_enabled=0))
    s.run([("style", 0), ("reset_adv",), ("frame", 0, 0)])
    nnf = _identity_nnf(w, h)
    rng = np.random.default_rng(94)
    nnf["x"] = rng.uniform(0, 30, (h, w)).astype(np.float32)
    nnf["y"] = rng.uniform(0, 30, (h, w)).astype(np.float32)
    o.load_offsets(0, nnf)
    o.sync_buffers(0, _ffi.PT_DONT_COPY_OFFSETS)
    s.run([("flip",), ("frame", 0, 1), ("advect", 1, 0)])
    return nnf, o.get_advected(0, 0, w, h)


def test_advect_zero_velocity_keeps_offsets():
    h, w = 20, 24
    z = np.zeros((h, w, 2), np.float32)
    with OracleContext() as o:
        nnf, (adv, base, out, res, ts) = _advect_setup(o, z, z)
    assert np.all(adv["hysteresis"] == 1) and np.all(ts == 1)
    assert adv["x"].tobytes() == nnf["x"].tobytes() and adv["y"].tobytes() == nnf["y"].tobytes()


def test_advect_integer_translation_shifts_offsets():
    h, w = 20, 24
    vf = np.zeros((h, w, 2), np.float32)
    vf[..., 0], vf[..., 1] = 3, 1
    vb = -vf
    with OracleContext() as o:
        nnf, (adv, base, out, res, ts) = _advect_setup(o, vf, vb)
    inner = adv[1:, 3:]
    assert np.all(inner["hysteresis"] == 1)
    assert inner["x"].tobytes() == nnf["x"][:-1, :-3].tobytes()
    assert inner["y"].tobytes() == nnf["y"][:-1, :-3].tobytes()
    assert np.all(adv[0]["hysteresis"] == 0) and np.all(ts[0] == -1)


def test_advect_inconsistent_motion_falls_back_to_random():
    h, w = 20, 24
    vb = np.zeros((h, w, 2), np.float32)
    vb[..., 0], vb[..., 1] = -3, -1
    with OracleContext() as o:
        nnf, (adv, base, out, res, ts) = _advect_setup(o, np.zeros((h, w, 2), np.float32), vb)
    assert np.all(adv["hysteresis"] == 0) and np.all(ts == -1)
    assert np.all(base == 0) and np.all(out == 0)


# ---------------------------------------------------------------- merges
def test_merge_randomly_extremes_and_residual_merge():
    h, w = 20, 24
    z = np.zeros((h, w, 2), np.float32)
    with OracleContext() as o:
        nnf, (adv, *_r) = _advect_setup(o, z, z)
        # after advection base == advected offsets; load a different field
        other = _identity_nnf(w, h)
        o.load_offsets(0, other)
        o.merge_randomly(0, True, 0.0)
        off, _, _ = o.get_output(0, w, h)
        assert off["x"].tobytes() == other["x"].astype(np.float32).tobytes()
        o.load_offsets(0, other)
        o.merge_randomly(0, True, 1.0)
        off, _, _ = o.get_output(0, w, h)
        assert off["x"].tobytes() == adv["x"].tobytes()
        # residual merge: advected residual 0 everywhere beats any positive residual
        o.load_offsets(0, other)
        o.load_residual(0, np.full((h, w), 5.0, np.float32))
        o.merge_using_residual(0, True)
        off, _, _ = o.get_output(0, w, h)
        adv_res = o.get_advected(0, 0, w, h)[3]
        pick_adv = adv_res < 5.0
        assert np.all(off["hysteresis"][pick_adv] == 1)
        assert off["x"][pick_adv].tobytes() == adv["x"][pick_adv].tobytes()


# ---------------------------------------------------------------- random init
def test_random_pyramid_ranges_and_alpha_rule():
    """rand() % (W_ex - 1) never picks the last column/row; fractional parts
    are rand()/RAND_MAX in [0, 1]; opaque pixels start on opaque exemplar
    texels when any exist (synthesisFrame.cpp:102-127)."""
    st = W.style_pair(40, 30, seed=101)
    st["output"][..., 3] = 0.1
    st["output"][10:20, 5:25, 3] = 1.0
    fr = W.still_frame(32, 32, seed=102)
    fr["color"][..., 3] = 1.0
    with OracleContext() as o:
        s = Session(o, st, [fr], o.default_params(residual_window_size=2, levels=1))
        s.run([("style", 0), ("frame", 0, 0), ("randomize",)])
        off, _, _ = s.state()
    ox, oy = np.floor(off["x"]).astype(int), np.floor(off["y"]).astype(int)
    fx, fy = off["x"] - np.floor(off["x"]), off["y"] - np.floor(off["y"])
    assert ox.min() >= 0 and ox.max() <= 39 and oy.min() >= 0 and oy.max() <= 29
    assert fx.min() >= 0 and fx.max() < 1 and fy.min() >= 0 and fy.max() < 1
    # texel actually chosen is the integer draw (x + rx with rx <= 1)
    draw_x = np.where(fx == 0, ox - 1, ox).clip(0)
    draw_y = np.where(fy == 0, oy - 1, oy).clip(0)
    opaque = (draw_x >= 5) & (draw_x < 25) & (draw_y >= 10) & (draw_y < 20)
    assert opaque.mean() > 0.99


# ---------------------------------------------------------------- IND schedule
def test_ind_first_frame_is_the_one_frame_schedule():
    """renderAnimationIND synthesizes frame 0 exactly like a one-frame FMBM_RM
    run (fresh buffers, random init at the coarsest level, upsampling from
    the same frame's pass 0); later frames run too and stay finite."""
    st = W.style_pair(40, 36, seed=61, guides=True)
    frames = W.animation(36, 32, 3, seed=62)
    with OracleContext() as o:
        p = o.default_params(residual_window_size=2, levels=3, temporal_coherence_weight=0.5)
        ind = o.synthesize(p, frames, st, scheme=_ffi.PT_S_IND, iterations=2, fixup_passes=0)
        one = o.synthesize(p, frames[:1], st, scheme=_ffi.PT_S_FMBM_RM, iterations=2, fixup_passes=0)
        ind2 = o.synthesize(p, frames, st, scheme=_ffi.PT_S_IND, iterations=2, fixup_passes=1)
    for a, b in zip(ind, one):
        assert a[0].tobytes() == b[0].tobytes()
    for a in ind2[1:]:
        assert np.isfinite(a).all()
    assert (ind2[2] >= 0).all()


def test_fbmo_single_level_single_frame_is_the_one_frame_schedule():
    """renderAnimationFBMO on one frame and one level: the forward and the
    backward pass run the same first-frame synthesis, and the force-valid
    merge of two equal fields returns it, so the output offsets are the
    one-frame FMBM_RM result.  (With more levels the two directed passes of
    a finer level differ through the upsample's untouched one-pixel ring,
    which holds what the previous pass left there.)"""
    st = W.style_pair(40, 36, seed=61, guides=True)
    frames = W.animation(36, 32, 1, seed=62)
    with OracleContext() as o:
        p = o.default_params(residual_window_size=2, levels=1, temporal_coherence_weight=0.5)
        a = o.synthesize(p, frames, st, scheme=_ffi.PT_S_FBMO, iterations=2, fixup_passes=0)
        b = o.synthesize(p, frames, st, scheme=_ffi.PT_S_FMBM_RM, iterations=2, fixup_passes=0)
    assert a[0].tobytes() == b[0].tobytes()


def test_unknown_schemes_are_rejected():
    st = W.style_pair(32, 32, seed=63)
    fr = W.still_frame(32, 32, seed=64)
    with OracleContext() as o:
        p = o.default_params(residual_window_size=2, levels=2)
        for scheme in (-1, 5):
            with pytest.raises(Exception):
                o.synthesize(p, [fr], st, scheme=scheme, iterations=1, fixup_passes=0)


@pytest.mark.parametrize("nframes,fixup", [(1, 0), (3, 1)])
def test_tctf_runs_and_keeps_offsets_inside_the_exemplar(nframes, fixup):
    """renderAnimationTCTF: every frame's offsets end within a texel of the
    exemplar and its energies are finite and non-negative (the schedule is
    pinned through its operators and the GPU parity; it has no closed
    form)."""
    st = W.style_pair(40, 36, seed=69, guides=True)
    frames = W.animation(36, 32, nframes, seed=70)
    with OracleContext() as o:
        p = o.default_params(residual_window_size=2, levels=2, temporal_coherence_weight=0.5)
        off, canvas, res = o.synthesize(p, frames, st, scheme=_ffi.PT_S_TCTF, iterations=2, fixup_passes=fixup)
    assert off.shape[0] == nframes
    # candidates are clamped to [0, size-1]; upsampled (2c+0.5) and averaged
    # offsets may lie up to a texel beyond, as in the other schedules
    assert (off["x"] >= 0).all() and (off["x"] <= 41).all() and (off["y"] >= 0).all() and (off["y"] <= 37).all()
    assert np.isfinite(canvas).all() and np.isfinite(res).all() and (res >= 0).all()


# ---------------------------------------------------------------- object IDs
def test_vote_downsample_known_answers():
    """downsampleVote (downsample.cpp:162-210) on hand-made 2x2 blocks: a
    3-1 and a 2-1-1 majority are the reference's answers; the 2-2 and the
    all-distinct ties go to the first ID in scan order (the reference's tie
    winner depends on QHash iteration order)."""
    ids = np.array([[5, 5, 1, 2],
                    [5, 9, 2, 3],
                    [4, 7, 8, 6],
                    [7, 4, 3, 1]], np.int32)
    with OracleContext() as o:
        p = o.pyramid(_ffi.PT_ELEM_I32, ids)
        assert p.levels == 2
        np.testing.assert_array_equal(p.download(0), ids)
        np.testing.assert_array_equal(p.download(1), [[5, 2], [4, 8]])
        q = o.pyramid(_ffi.PT_ELEM_I32, None, default=3.0, shape=(8, 4))
        assert (q.download(1) == 3).all()


def test_object_id_term_point_fetch_and_weight():
    """The term is added at the centre tap only, weight * g_c / sum(g) of the
    patch distance, when the image ID at the pixel differs from the style ID
    at the candidate's centre texel floor(x + 0.5): a candidate at x = 5.49
    reads texel 5, one at x = 5.5 texel 6."""
    st = W.style_pair(16, 12, seed=47)
    fr = W.still_frame(12, 10, seed=48)
    st["object_ids"] = np.zeros((12, 16), np.int32)
    st["object_ids"][:, 5] = 1
    st["object_ids"][:, 6] = 2
    fr["object_id"] = np.ones((10, 12), np.int32)
    r, wt = 2, 0.75
    with OracleContext() as o:
        g = None
        d = {}
        for w in (0.0, wt):
            s = Session(o, st, [fr], o.default_params(residual_window_size=r, levels=1, style_obj_id_weight=w))
            s.run([("style", 0), ("frame", 0, 0), ("randomize",)])
            g = o.gaussian_weights(r)
            for cx in (5.49, 5.5):
                cand = np.zeros((), _ffi.PATCHXF_DTYPE)
                cand["x"], cand["y"], cand["scale_u"], cand["scale_v"] = cx, 4.0, 1.0, 1.0
                d[w, cx] = o.patch_distance(6, 5, 0, cand)
        assert d[wt, 5.49] == d[0.0, 5.49]
        exp = wt * float(g[r]) ** 2 / float(np.sum(g)) ** 2
        assert abs((d[wt, 5.5] - d[0.0, 5.5]) - exp) <= 1e-6 * exp + 1e-6 * d[0.0, 5.5]


@pytest.mark.parametrize("r,level", [(2, 0), (3, 1)])
def test_patch_distance_with_object_ids_matches_float64(r, level):
    st = W.style_pair(24, 20, seed=49, guides=True)
    fr = W.still_frame(22, 18, seed=50, guides=True)
    st["object_ids"] = W.object_ids(49, 30, 24, 20)
    fr["object_id"] = W.object_ids(50, 31, 22, 18)
    rng = np.random.default_rng(51)
    with OracleContext() as o:
        prm = o.default_params(residual_window_size=r, levels=2, style_obj_id_weight=0.9, dist_trans_weight=0.4)
        s = Session(o, st, [fr], prm)
        s.run([("style", level), ("frame", level, 0), ("randomize",)])
        off, canvas, _ = s.state()
        exb, exo = s.sb.download(level), s.so.download(level)
        exdt = s.sdt.download(level)
        target = s.frames[0]["color"].download(level)
        imgdt = s.frames[0]["dt"].download(level)
        ids = (s.frames[0]["id"].download(level), s.sid.download(level))
        g = o.gaussian_weights(r)
        lh, lw = target.shape[:2]
        hits = 0
        for x, y in [(0, 0), (lw - 1, lh - 1)] + [(int(a), int(b)) for a, b in
                                                  zip(rng.integers(0, lw, 14), rng.integers(0, lh, 14))]:
            cand = np.zeros((), _ffi.PATCHXF_DTYPE)
            cand["x"] = rng.uniform(0, exb.shape[1])
            cand["y"] = rng.uniform(0, exb.shape[0])
            cand["theta"] = rng.uniform(-1.5, 1.5)
            cand["scale_u"] = cand["scale_v"] = 1.0
            got = o.patch_distance(x, y, level, cand)
            exp = _pde_double(x, y, cand, target, canvas, exb, exo, exdt, imgdt, g, prm, ids=ids)
            assert abs(got - exp) <= 3e-6 * max(exp, 1e-3), (x, y, got, exp)
            hits += _pde_double(x, y, cand, target, canvas, exb, exo, exdt, imgdt, g, prm, ids=ids) != \
                _pde_double(x, y, cand, target, canvas, exb, exo, exdt, imgdt, g,
                            o.default_params(residual_window_size=r, levels=2, dist_trans_weight=0.4))
        assert hits >= 3  # the term fired on some of the candidates


def test_object_ids_absent_or_equal_change_nothing():
    """With the weight on but no ID images (all 0), or with equal IDs
    everywhere, the schedule gives the weight-0 result bit for bit."""
    st = W.style_pair(32, 32, seed=52, guides=True)
    frames = W.animation(24, 20, 2, seed=53)
    kw = dict(residual_window_size=2, levels=2, temporal_coherence_weight=0.5)
    with OracleContext() as o:
        base = o.synthesize(o.default_params(**kw), frames, st, iterations=2)
        off = o.synthesize(o.default_params(style_obj_id_weight=0.8, **kw), frames, st, iterations=2)
        same = [dict(f, object_id=np.full((20, 24), 11, np.int32)) for f in frames]
        st2 = dict(st, object_ids=np.full((32, 32), 11, np.int32))
        eq = o.synthesize(o.default_params(style_obj_id_weight=0.8, **kw), same, st2, iterations=2)
        diff = o.synthesize(o.default_params(style_obj_id_weight=0.8, **kw),
                            [dict(f, object_id=W.object_ids(53, 40 + i, 24, 20)) for i, f in enumerate(frames)],
                            dict(st, object_ids=W.object_ids(52, 41, 32, 32)), iterations=2)
    for a, b, c in zip(base, off, eq):
        assert a.tobytes() == b.tobytes() == c.tobytes()
    assert base[0].tobytes() != diff[0].tobytes()


# ---------------------------------------------------------------- luminance shift
def test_luminance_delta_matches_float64():
    """Color4::toLuminance of a colour difference (types.h:87-89): the float
    differences, then 0.2126 r + 0.7152 g + 0.0722 b in double, to float."""
    lib = load_oracle()
    rng = np.random.default_rng(61)
    for _ in range(2000):
        c = rng.uniform(-0.5, 1.5, 3).astype(np.float32)
        p = rng.uniform(-0.5, 1.5, 3).astype(np.float32)
        d = (c - p).astype(np.float32)
        want = np.float32((0.2126 * float(d[0]) + 0.7152 * float(d[1])) + 0.0722 * float(d[2]))
        got = lib.ptref_luminance_delta(*(float(v) for v in c), *(float(v) for v in p))
        assert np.float32(got) == want
    assert lib.ptref_luminance_delta(0.3, 0.4, 0.5, 0.3, 0.4, 0.5) == 0.0


def _lum_frames():
    f0 = W.still_frame(24, 20, seed=62)
    f1 = dict(f0, color=f0["color"].copy())
    f1["color"][..., :3] += np.float32(0.1)
    return f0, f1


def test_luminance_shift_advection_and_canvas():
    """With lumShiftEnable, an offset advected along a zero velocity field
    carries the source's shift plus the luminance change of its pixel
    between the frames (texSynth_kernel.cu:1626-1631), and the canvas adds
    the shift to R, G and B (toColor4(c, lum), :272)."""
    st = W.style_pair(32, 28, seed=63)
    f0, f1 = _lum_frames()
    ops = [("style", 0), ("reset_adv",), ("frame", 0, 0), ("randomize",), ("sync", 0), ("flip",), ("frame", 0, 1),
           ("params", dict(direction=_ffi.PT_FORWARD, frame_number=1)), ("advect", 1, 0)]
    out = {}
    for on in (0, 1):
        with OracleContext() as o:
            s = Session(o, st, [f0, f1], o.default_params(residual_window_size=2, levels=1, lum_shift_enable=on))
            out[on] = s.run(ops)[-1][1]
    off1, can1, _ = out[1]
    off0, can0, _ = out[0]
    lib = load_oracle()
    valid = off1["hysteresis"] > 0
    assert valid.mean() > 0.9
    for y, x in zip(*np.nonzero(valid)):
        c, p = f1["color"][y, x], f0["color"][y, x]
        assert off1["lum_y"][y, x] == np.float32(lib.ptref_luminance_delta(*map(float, c[:3]), *map(float, p[:3])))
    assert (off1["lum_y"][~valid] == 0).all() and (off0["lum_y"] == 0).all()
    assert abs(off1["lum_y"][valid].mean() - 0.1) < 1e-6  # the coefficients sum to 1
    for k in off0.dtype.names:
        if k != "lum_y":
            assert off0[k].tobytes() == off1[k].tobytes(), k
    np.testing.assert_array_equal(can1[..., :3], (can0[..., :3] + off1["lum_y"][..., None]).astype(np.float32))
    np.testing.assert_array_equal(can1[..., 3], can0[..., 3])


def test_luminance_shift_kept_or_reset_by_propagation():
    """A pixel that keeps its offset keeps its shift; one that takes a
    propagated candidate gets Luminance() (0)."""
    st = W.style_pair(32, 28, seed=64)
    f0, _ = _lum_frames()
    with OracleContext() as o:
        s = Session(o, st, [f0], o.default_params(residual_window_size=2, levels=1, lum_shift_enable=1))
        s.run([("style", 0), ("frame", 0, 0), ("randomize",)])
        off, _, _ = s.state(0)
        off["lum_y"] = np.float32(0.25)
        o.load_offsets(0, off)
        o.sync_buffers(0, _ffi.PT_DONT_COPY_OFFSETS)
        o.propagate(0, 1)
        o.sync_buffers(0)
        new, _, _ = s.state(0)
        moved = (new["x"] != off["x"]) | (new["y"] != off["y"]) | (new["theta"] != off["theta"])
        assert 0.05 < moved.mean() < 0.95
        assert (new["lum_y"][moved] == 0).all() and (new["lum_y"][~moved] == np.float32(0.25)).all()
        bad = off.copy()
        bad["lum_a"][0, 0] = 0.5
        with pytest.raises(Exception):
            o.load_offsets(0, bad)


# ---------------------------------------------------------------- several styles
@pytest.mark.parametrize("styles", [2, 3])
def test_patch_distance_wraps_at_style_height_times_styles(styles):
    """With numStyles > 1 the tap coordinates wrap at the offsets histogram's
    height, styleHeight x numStyles (workingBuffers.cpp:57,86,
    texSynth_kernel.cu:595-600): taps below the exemplar clamp to its last row
    instead of wrapping to the top."""
    st = W.style_pair(24, 20, seed=53)
    fr = W.still_frame(22, 18, seed=54)
    rng = np.random.default_rng(55)
    with OracleContext() as o:
        vals = {}
        for n in (1, styles):
            prm = o.default_params(residual_window_size=3, levels=1, num_styles=n)
            s = Session(o, st, [fr], prm)
            s.run([("style", 0), ("frame", 0, 0), ("randomize",)])
            off, canvas, _ = s.state()
            exb, exo = s.sb.download(0), s.so.download(0)
            target = s.frames[0]["color"].download(0)
            g = o.gaussian_weights(3)
            got = []
            for k in range(12):
                cand = np.zeros((), _ffi.PATCHXF_DTYPE)
                cand["x"] = rng.uniform(4, 20)
                cand["y"] = rng.uniform(17.5, 19.5)  # taps cross the bottom edge
                cand["theta"] = rng.uniform(-0.3, 0.3)
                cand["scale_u"] = cand["scale_v"] = 1.0
                x, y = int(rng.integers(3, 19)), int(rng.integers(3, 15))
                v = o.patch_distance(x, y, 0, cand)
                exp = _pde_double(x, y, cand, target, canvas, exb, exo, None, None, g, prm, styles=n)
                assert abs(v - exp) <= 3e-6 * max(exp, 1e-3), (n, x, y, v, exp)
                got.append(v)
            vals[n] = got
            rng = np.random.default_rng(55)
        assert vals[1] != vals[styles]


def test_style_layers_above_zero_are_accepted():
    with OracleContext() as o:
        st = W.style_pair(24, 20, seed=56)
        fr = W.still_frame(22, 18, seed=57)
        s = Session(o, st, [fr], o.default_params(residual_window_size=2, levels=1, num_styles=3))
        o.initialize_exemplars(24, 20, 3)
        o.upload_image(0, "EXEMPLAR_BASE", s.sb, style_layer=2)
        with pytest.raises(Exception):
            o.upload_image(0, "EXEMPLAR_BASE", s.sb, style_layer=4)
        with pytest.raises(Exception):
            o.setup(o.default_params(num_styles=5))


# ---------------------------------------------------------------- first pass direction
def test_backward_first_pass_starts_from_the_last_frame():
    """firstPassDirection 1 (synthesisProcessor.cpp:707): pass 0 runs
    backwards, so the last frame is synthesized first, as a first frame (random
    initialisation, no advection): with one level (whose random pyramid does
    not depend on the frame number) its result is the one-frame synthesis of
    that frame, as frame 0's is with the forward first pass."""
    st = W.style_pair(40, 36, seed=58, guides=True)
    frames = W.animation(32, 28, 3, seed=59)
    kw = dict(residual_window_size=2, levels=1, temporal_coherence_weight=0.5)
    with OracleContext() as o:
        out = {d: o.synthesize(o.default_params(first_pass_direction=d, **kw), frames, st,
                               scheme=_ffi.PT_S_FMBM_RM, iterations=2, fixup_passes=0) for d in (0, 1)}
        alone = {f: o.synthesize(o.default_params(**kw), [frames[f]], st, scheme=_ffi.PT_S_FMBM_RM, iterations=2,
                                 fixup_passes=0) for f in (0, 2)}
        for a, b in zip(out[1], alone[2]):
            assert a[2].tobytes() == b[0].tobytes()
        for a, b in zip(out[0], alone[0]):
            assert a[0].tobytes() == b[0].tobytes()
        assert out[1][0][2].tobytes() != out[0][0][2].tobytes()
        with pytest.raises(Exception):
            o.synthesize(o.default_params(first_pass_direction=1, **kw), frames, st, scheme=_ffi.PT_S_TCTF)
        with pytest.raises(Exception):
            o.setup(o.default_params(first_pass_direction=2))


# ---------------------------------------------------------------- maxLevel
def test_max_level_upsamples_the_coarse_result_without_synthesis():
    """maxLevel 1 (renderAnimationFMBM, synthesisProcessor.cpp:849-872): level
    1 is synthesized, then every frame's result is upsampled to level 0 and
    its canvas synced, with no image reload and no search, and the output is
    that pass 0.  One frame: the same as the operator sequence replayed on a
    fresh context (a used one holds stale level-0-sized buffers that the
    coarse level reads).  Also what the upsample alone leaves: hysteresis 0
    on every pixel it writes (the full schedule's search does not), offsets
    within the upsampled exemplar and an output unlike the full schedule's."""
    st = W.style_pair(40, 36, seed=65, guides=True)
    fr = W.still_frame(36, 32, seed=66, guides=True)
    kw = dict(residual_window_size=2, levels=2, orientation_enabled=1, dist_trans_weight=0.5)
    with OracleContext() as o:
        got = o.synthesize(o.default_params(max_level=1, **kw), [fr], st, scheme=_ffi.PT_S_FMBM_RM, iterations=2,
                           fixup_passes=0)
        full = o.synthesize(o.default_params(**kw), [fr], st, scheme=_ffi.PT_S_FMBM_RM, iterations=2,
                            fixup_passes=0)
        off = got[0][0]
        assert (off["hysteresis"][1:-1, 1:-1] == 0).all()
        assert (full[0][0]["hysteresis"][1:-1, 1:-1] != 0).any()
        assert off.tobytes() != full[0][0].tobytes()
        inner = off[1:-1, 1:-1]
        # 2c + 0.5 of coarse offsets that lie up to a texel beyond the 20x18 level-1 exemplar
        assert (inner["x"] >= 0.5).all() and (inner["x"] <= 41.5).all()
        assert (inner["y"] >= 0.5).all() and (inner["y"] <= 37.5).all()
        assert np.isfinite(got[1]).all() and np.isfinite(got[2]).all()
    with OracleContext() as o:
        s = Session(o, st, [fr], o.default_params(**kw))
        sweep = [("propagate", j) for j in (128, 64, 32, 16, 8, 4, 2, 1, 1)]
        s.run([("style", 1), ("reset_adv",), ("frame", 1, 0), ("randomize",), ("sync", 0)] + sweep +
              [("scatter",)] + sweep)
        off1, _, _ = s.state(1)
        s.load_style(0)
        o.load_offsets(1, off1)
        o.upsample(0)
        o.sync_buffers(0, _ffi.PT_DONT_COPY_OFFSETS | _ffi.PT_DONT_UPDATE_RESIDUAL)
        for a, b in zip(got, s.state(0)):
            assert a[0].tobytes() == b.tobytes()
    with OracleContext() as o:
        with pytest.raises(Exception):
            o.synthesize(o.default_params(max_level=1, **kw), [fr], st, scheme=_ffi.PT_S_IND)
        with pytest.raises(Exception):
            o.setup(o.default_params(max_level=2, **kw))


def test_offsets_histogram_term_is_zero_as_in_the_reference():
    """histogramTerm (texSynth_kernel.cu:486-494) reads a histogram that is
    never accumulated (:258) and is reset and blurred from zeros
    (:1250-1276): any slope with a non-negative threshold changes no bit; a
    negative layer-0 threshold with a nonzero slope, which would add a
    constant, is refused; thresholds of layers the candidates never use
    (styles above 0 without keyframes) do not matter."""
    st = W.style_pair(40, 36, seed=91, guides=True)
    fr = W.still_frame(36, 32, seed=92, guides=True)
    kw = dict(residual_window_size=2, levels=2, dist_trans_weight=0.5)
    runs = []
    for slope, thr in ((0.0, 0.0), (2.0, 0.3), (7.5, 0.0), (0.0, -0.5)):
        with OracleContext() as o:
            p = o.default_params(**kw)
            p.offsets_histogram_slope[0] = slope
            p.offsets_histogram_threshold[0] = thr
            p.offsets_histogram_slope[3] = 1.0
            p.offsets_histogram_threshold[3] = -4.0
            runs.append(o.synthesize(p, [fr], st, scheme=_ffi.PT_S_FMBM_RM, iterations=2, fixup_passes=0))
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            assert a.tobytes() == b.tobytes()
    with OracleContext() as o:
        p = o.default_params(**kw)
        p.offsets_histogram_slope[0] = 2.0
        p.offsets_histogram_threshold[0] = -0.5
        with pytest.raises(Exception):
            o.setup(p)
