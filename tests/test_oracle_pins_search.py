"""Pins of the oracle's search operators against an independent float64
statement of the reference kernels, so a plausible mistake in the
restatement (a sign of (x - tx), a direction of the angle fold, a wrong
neighbour set, a dropped temporal term, a bilinear read where the reference
reads a texel) fails here:

  timeDerivativeTerms / hysteresisTerm   texSynth_kernel.cu:438-469, 506-535
  screenToExemplarAngle                  :425-436
  calculatePropagateOffset + propagate   :956-982, 1112-1219
  scatter_kernel (truncated RNG)         :865-952
  averageOffset                          :841-862
  upsampleBilinearKernel nearest branch  :1403-1417, 1453-1510
  culling                                :176-205, 678-681, 880-883, 1131-1135, 1592-1608

The float64 statement here reads the oracle's textures and buffers back and
recomputes each quantity from the reference's definitions; tolerances cover
only the float32 rounding of the oracle."""
import numpy as np
import pytest

from painttween_amd import _ffi
from painttween_amd import workloads as W
from tests.harness import Session
from tests.oracle_ctx import OracleContext

E1, E2 = 1, 2
HALF_PI = np.pi / 2


# ---------------------------------------------------------------- float64 statement
def _bilinear(tex, x, y):
    """Clamp-addressed bilinear fetch of a (h, w[, c]) texture at texel-space
    (x, y) (PYRAMID_FETCH's +0.5 texel centre), float64, vectorised."""
    t = tex.astype(np.float64)
    if t.ndim == 2:
        t = t[..., None]
    h, w = t.shape[:2]
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    fx, fy = np.floor(x), np.floor(y)
    a, b = (x - fx)[..., None], (y - fy)[..., None]
    x0, y0 = fx.astype(np.int64), fy.astype(np.int64)
    xa, xb = np.clip(x0, 0, w - 1), np.clip(x0 + 1, 0, w - 1)
    ya, yb = np.clip(y0, 0, h - 1), np.clip(y0 + 1, 0, h - 1)
    return (1 - a) * (1 - b) * t[ya, xa] + a * (1 - b) * t[ya, xb] + (1 - a) * b * t[yb, xa] + a * b * t[yb, xb]


def _point(tex, x, y):
    """Point-filtered fetch: texel floor(x + 0.5), clamped."""
    h, w = tex.shape[:2]
    return tex[min(max(int(np.floor(y + 0.5)), 0), h - 1), min(max(int(np.floor(x + 0.5)), 0), w - 1)]


def _fold(theta):
    """screenToExemplarAngle's fold of a float32 angle difference into
    [-pi/2, pi/2] with the double constants M_PI_2 / M_PI."""
    t = float(theta)
    if t > HALF_PI:
        return t - np.pi
    if t < -HALF_PI:
        return np.pi + t
    return t


class State:
    """The oracle's textures and buffers at one level, read back."""

    def __init__(self, o, s, level, frame=None):
        self.o, self.s, self.level, self.prm = o, s, level, s.params
        self.lw, self.lh = s.w >> level, s.h >> level
        self.off, self.canvas, self.res = o.get_output(level, self.lw, self.lh)
        self.exb, self.exo = s.sb.download(level), s.so.download(level)
        self.exdt, self.exor = s.sdt.download(level), s.sor.download(level)
        self.ehh = self.exb.shape[0] * self.prm.num_styles
        f = s.frames[frame] if frame is not None else None
        if f is not None:
            self.target = f["color"].download(level)
            self.imgdt = f["dt"].download(level)
            self.imgor = f["orient"].download(level)
            self.scale = f["scale"].download(level)
            self.velb, self.velf = f["velb"].download(level), f["velf"].download(level)
        self.g = np.array(o.gaussian_weights(self.prm.residual_window_size), np.float64)
        self.adv = None

    def load_advected(self):
        self.adv = [self.o.get_advected(self.level, k, self.lw, self.lh) for k in (0, 1)]

    # screenToExemplarAngle at image (x, y), exemplar (ex, ey)
    def angle(self, x, y, ex, ey):
        if not self.prm.orientation_enabled:
            return 0.0
        d = np.float32(_point(self.exor, ex, ey)) - np.float32(_point(self.imgor, x, y))
        return _fold(d)

    def basis(self, th, su, sv):
        p = self.prm
        su, sv = p.residual_window_scale_u * su, p.residual_window_scale_v * sv
        return (np.cos(th) * su, np.sin(th) * su), (-np.sin(th) * sv, np.cos(th) * sv)

    def hysteresis(self, x, y):
        """hysteresisTerm (:506-535): velocities at the 2x2 texels from (x, y),
        B when the pass runs forward, F otherwise; the weight if both spreads
        are below maxDistortion / 2^level."""
        p = self.prm
        v = self.velb if p.direction == _ffi.PT_FORWARD else self.velf
        h, w = v.shape[:2]
        q = np.array([v[min(y + dy, h - 1), min(x + dx, w - 1)] for dx, dy in ((0, 0), (1, 0), (0, 1), (1, 1))],
                     np.float64)
        md = p.max_distortion / 2.0 ** self.level
        if q[:, 0].max() - q[:, 0].min() < md and q[:, 1].max() - q[:, 1].min() < md:
            return float(p.hysteresis_weight)
        return 0.0

    def pde(self, x, y, cand, time=True):
        """calculateResidual (:537-659), every term the configuration turns on
        (temporal terms and hysteresis from the advected buffers when loaded)."""
        p = self.prm
        r = p.residual_window_size
        ii, jj = np.meshgrid(np.arange(-r, r + 1), np.arange(-r, r + 1), indexing="ij")
        ii, jj = ii.ravel(), jj.ravel()  # i (x) outer, j (y) inner
        inb = (x + ii >= 0) & (x + ii < self.lw) & (y + jj >= 0) & (y + jj < self.lh)
        ii, jj = ii[inb], jj[inb]
        (b0x, b0y), (b1x, b1y) = self.basis(float(cand["theta"]), float(cand["scale_u"]), float(cand["scale_v"]))
        tx = float(cand["x"]) + ii * b0x + jj * b1x
        ty = float(cand["y"]) + ii * b0y + jj * b1y
        hw, hh = self.exb.shape[1], self.ehh
        tx = np.where(tx >= hw, tx - hw, np.where(tx < 0, tx + hw, tx))
        ty = np.where(ty >= hh, ty - hh, np.where(ty < 0, ty + hh, ty))
        tb, to = _bilinear(self.exb, tx, ty), _bilinear(self.exo, tx, ty)
        base = self.target[y + jj, x + ii].astype(np.float64)
        canv = self.canvas[y + jj, x + ii].astype(np.float64)
        aw = np.array([1, 1, 1, p.alpha_weight], np.float64)
        res = ((base - tb) ** 2 * aw).sum(1) * p.input_analogy_weight
        centre = (ii == 0) & (jj == 0)
        res += np.where(centre, 0.0, ((canv - to) ** 2 * aw).sum(1) * p.canvas_output_weight)
        if p.dist_trans_weight > 0:
            k = np.nonzero(centre)[0][0]
            de = _bilinear(self.exdt, tx[k], ty[k])[0] / float(cand["scale_v"])
            res[k] += min(abs(de - float(self.imgdt[y, x])), 10.0) * p.dist_trans_weight
        num = 0.0
        if time and self.adv is not None:
            any_w = p.time_derivative_input_weight > 0 or p.time_derivative_output_weight > 0 or \
                p.temporal_coherence_weight > 0
            fwd_on = p.direction in (_ffi.PT_FORWARD, _ffi.PT_BIDIRECTIONAL)
            bwd_on = p.direction in (_ffi.PT_BACKWARD, _ffi.PT_BIDIRECTIONAL)
            if self.adv[0][0]["hysteresis"][y, x] > 0:
                num -= self.hysteresis(x, y)
            for k, on in ((0, fwd_on), (1, bwd_on)):
                aoff, abase, aout = self.adv[k][0], self.adv[k][1], self.adv[k][2]
                if not (on and any_w and aoff["hysteresis"][y, x] > 0):
                    continue
                valid = aoff["hysteresis"][y + jj, x + ii] > 0
                prev = aoff[y, x]
                th = self.angle(x, y, float(prev["x"]), float(prev["y"]))
                (c0x, c0y), (c1x, c1y) = self.basis(th, float(prev["scale_u"]), float(prev["scale_v"]))
                px = float(prev["x"]) + ii * c0x + jj * c1x  # no wrap for the previous offset
                py = float(prev["y"]) + ii * c0y + jj * c1y
                peb, peo = _bilinear(self.exb, px, py), _bilinear(self.exo, px, py)
                pb = abase[y + jj, x + ii].astype(np.float64)
                po = aout[y + jj, x + ii].astype(np.float64)
                t = (((base - pb) - (tb - peb)) ** 2).sum(1) * p.time_derivative_input_weight
                t += (((canv - po) - (to - peo)) ** 2).sum(1) * p.time_derivative_output_weight
                t += ((po - to) ** 2).sum(1) * p.temporal_coherence_weight
                res += np.where(valid, t, 0.0)
        wgt = self.g[ii + r] * self.g[jj + r]
        return (num + (res * wgt).sum()) / wgt.sum()

    def test_offset(self, x, y, cand, image_alpha):
        """test_offset (:695-721): exemplar culling at the truncated texel,
        the alpha rule, then the patch distance."""
        p = self.prm
        h, w = self.exdt.shape
        ex = min(max(int(float(cand["x"])), 0), w - 1)  # int parameter: truncation
        ey = min(max(int(float(cand["y"])), 0), h - 1)
        if float(self.exdt[ey, ex]) > p.dist_trans_cull * 1.25:
            return 1e31
        if not p.transparency_ok:
            ea = _bilinear(self.exo, float(cand["x"]), float(cand["y"]))[3]
            if ea < 0.99 and image_alpha > 0.99:
                return 1e31
        return self.pde(x, y, cand)

    def image_culled(self, x, y):
        return float(self.imgdt[y, x]) > self.prm.dist_trans_cull


def _cand(x, y, th, hy=0.0, su=1.0, sv=1.0):
    c = np.zeros((), _ffi.PATCHXF_DTYPE)
    c["x"], c["y"], c["theta"], c["hysteresis"], c["scale_u"], c["scale_v"] = x, y, th, hy, su, sv
    return c


def _select(start_res, cands, ress, tie=1e-5):
    """The reference's strict in-order selection in float64 plus the set of
    outcomes a float32 evaluation may also reach through a near-tie."""
    best, win = start_res, -1
    for k, rk in enumerate(ress):
        if rk < best:
            best, win = rk, k
    ok = {k for k, rk in enumerate(ress) if rk <= best * (1 + tie) + 1e-12}
    if start_res <= best * (1 + tie) + 1e-12:
        ok.add(-1)
    ok.add(win)
    return win, ok


def _same_xy(a, c, tol=2e-3):
    return abs(float(a["x"]) - float(c["x"])) <= tol and abs(float(a["y"]) - float(c["y"])) <= tol


# ---------------------------------------------------------------- angle fold
def test_screen_to_exemplar_angle_fold_edges():
    """theta = exemplar orientation (point) - image orientation (point),
    folded by pi when it passes +-pi/2 as a double; float32(pi/2) is just above
    pi/2 and folds, the float below it does not; orientation off gives 0."""
    f32 = np.float32
    hp = f32(np.pi / 2)
    below = np.nextafter(hp, f32(0))
    vals = [hp, below, -hp, -below, f32(1.0), f32(-1.0), f32(3.0), f32(-3.0), f32(2.0), f32(-2.0), f32(0.0)]
    ex_or = np.zeros((4, 16), np.float32)
    ex_or[1, :len(vals)] = vals
    img = {"color": W.sin_noise(1, 2, 16, 16)}
    st = {"base": W.sin_noise(1, 0, 16, 4), "output": W.sin_noise(1, 1, 16, 4)}
    with OracleContext() as o:
        s = Session(o, st, [img], o.default_params(residual_window_size=1, levels=1))
        s.run([("style", 0), ("frame", 0, 0)])
        o.upload_image(0, "EXEMPLAR_ORIENTATION", o.pyramid(E1, ex_or))
        for io in (0.0, 0.5, -1.25):
            o.upload_image(0, "INPUT_ORIENTATION", o.pyramid(E1, np.full((16, 16), io, np.float32)))
            for k, v in enumerate(vals):
                # (k + 0.4, 1.3) point-samples texel (k, 1); image pixel (3, 5)
                got = o.screen_to_exemplar_angle(3.0, 5.0, k + 0.4, 1.3)
                d = float(np.float32(v) - np.float32(io))  # the float difference, compared as a double
                exp = np.float32(d - np.pi if d > np.pi / 2 else (np.pi + d if d < -np.pi / 2 else d))
                assert got == exp, (io, v, got, exp)
        assert o.screen_to_exemplar_angle(3.0, 5.0, 0.4, 1.3) == np.float32(float(hp - np.float32(-1.25)) - np.pi)
        # exact edge values: float32(pi/2) folds to float32(pi/2) - pi, the float below stays
        o.upload_image(0, "INPUT_ORIENTATION", o.pyramid(E1, np.zeros((16, 16), np.float32)))
        assert o.screen_to_exemplar_angle(3.0, 5.0, 0.0, 1.0) == np.float32(float(hp) - np.pi)
        assert o.screen_to_exemplar_angle(3.0, 5.0, 1.0, 1.0) == below
        assert o.screen_to_exemplar_angle(3.0, 5.0, 2.0, 1.0) == np.float32(np.pi - float(hp))
        assert o.screen_to_exemplar_angle(3.0, 5.0, 3.0, 1.0) == -below
        p = s.params
        p.orientation_enabled = 0
        o.upload_parameters(p)
        assert o.screen_to_exemplar_angle(3.0, 5.0, 6.0, 1.0) == 0.0


# ---------------------------------------------------------------- temporal terms
def _temporal_state(o, r, direction, hyst=0.25, max_dist=0.6, level=0):
    """Frame 1 current with frame 0 advected forward into the F set and
    frame 2 advected backward into the B set (the order a backward pass
    loads them), both mostly valid."""
    st = W.style_pair(30, 28, seed=301, guides=True)
    frames = W.animation(26, 22, 3, seed=302)
    prm = o.default_params(residual_window_size=r, levels=level + 1, time_derivative_input_weight=0.7,
                           time_derivative_output_weight=1.3, temporal_coherence_weight=0.5, hysteresis_weight=hyst,
                           max_distortion=max_dist, dist_trans_weight=0.4, residual_window_scale_u=1.1,
                           residual_window_scale_v=0.9)
    s = Session(o, st, frames, prm)
    s.run([("style", level), ("reset_adv",), ("frame", level, 0), ("randomize",), ("flip",), ("frame", level, 1),
           ("advect", 1), ("flip",), ("frame", level, 2), ("flip",), ("frame", level, 1), ("advect", 0),
           ("params", {"direction": direction})])
    S = State(o, s, level, frame=1)
    S.load_advected()
    return s, S


@pytest.mark.parametrize("direction", [_ffi.PT_FORWARD, _ffi.PT_BACKWARD, _ffi.PT_BIDIRECTIONAL])
@pytest.mark.parametrize("r", [2, 3])
def test_temporal_terms_and_hysteresis_match_float64(direction, r):
    rng = np.random.default_rng(303 + direction + 10 * r)
    with OracleContext() as o:
        s, S = _temporal_state(o, r, direction)
        fv, bv = S.adv[0][0]["hysteresis"] > 0, S.adv[1][0]["hysteresis"] > 0
        assert fv.mean() > 0.5 and bv.mean() > 0.5, (fv.mean(), bv.mean())
        hy = [S.hysteresis(x, y) for y in range(1, S.lh - 1) for x in range(1, S.lw - 1)]
        assert 0 < sum(h > 0 for h in hy) < len(hy)  # the distortion test splits the pixels
        n_time = 0
        for x, y in [(0, 0), (S.lw - 1, S.lh - 1)] + list(zip(rng.integers(0, S.lw, 30), rng.integers(0, S.lh, 30))):
            x, y = int(x), int(y)
            for _ in range(3):
                cand = _cand(rng.uniform(1, 28), rng.uniform(1, 26), rng.uniform(-1.5, 1.5),
                             su=rng.uniform(0.8, 1.2), sv=rng.uniform(0.8, 1.2))
                got = o.patch_distance(x, y, 0, cand)
                exp = S.pde(x, y, cand)
                base = S.pde(x, y, cand, time=False)
                n_time += abs(exp - base) > 1e-3 * abs(base)
                assert abs(got - exp) <= 3e-6 * (abs(exp) + 1.0), (x, y, got, exp, base)
        assert n_time > 60  # the temporal terms changed most of the energies


def test_temporal_terms_each_weight_and_direction_matter():
    """Turning off one temporal weight or switching the direction changes the
    oracle's energy exactly as the float64 statement does."""
    rng = np.random.default_rng(310)
    with OracleContext() as o:
        s, S = _temporal_state(o, 2, _ffi.PT_BIDIRECTIONAL)
        cands = [(int(rng.integers(2, 24)), int(rng.integers(2, 20)),
                  _cand(rng.uniform(2, 27), rng.uniform(2, 25), rng.uniform(-1, 1))) for _ in range(25)]
        for change in ({"time_derivative_input_weight": 0.0}, {"time_derivative_output_weight": 0.0},
                       {"temporal_coherence_weight": 0.0}, {"hysteresis_weight": 0.0},
                       {"direction": _ffi.PT_FORWARD}, {"direction": _ffi.PT_BACKWARD}):
            before = [o.patch_distance(x, y, 0, c) for x, y, c in cands]
            s.run([("params", change)])
            after = [o.patch_distance(x, y, 0, c) for x, y, c in cands]
            for (x, y, c), g in zip(cands, after):
                exp = S.pde(x, y, c)
                assert abs(g - exp) <= 3e-6 * (abs(exp) + 1.0), (change, x, y, g, exp)
            assert sum(abs(a - b) > 1e-6 for a, b in zip(before, after)) >= 5, change
            s.run([("params", {"time_derivative_input_weight": 0.7, "time_derivative_output_weight": 1.3,
                               "temporal_coherence_weight": 0.5, "hysteresis_weight": 0.25,
                               "direction": _ffi.PT_BIDIRECTIONAL})])


# ---------------------------------------------------------------- propagate
def _search_session(o, level, dt_cull=False, transparency_ok=0, **kw):
    st = W.style_pair(30, 28, seed=321, guides=True)
    fr = W.still_frame(24, 20, seed=322, guides=True)
    fr["scale"] = np.stack([0.8 + 0.45 * W.smooth_noise(323, 0, 24, 20),
                            0.8 + 0.45 * W.smooth_noise(323, 1, 24, 20)], -1).astype(np.float32)
    if dt_cull:
        fr["dist_trans"] = W.culling_dist_trans(324, 0, 24, 20)
        st["dist_trans"] = W.culling_dist_trans(324, 1, 30, 28, lo=40.0, hi=80.0)
    prm = o.default_params(residual_window_size=kw.pop("r", 2), levels=2, dist_trans_weight=0.5,
                           transparency_ok=transparency_ok, **kw)
    s = Session(o, st, [fr], prm)
    ops = [("style", 0), ("frame", 0, 0), ("randomize",)]
    if level == 1:
        ops += [("style", 1), ("frame", 1, 0), ("randomize",)]
    s.run(ops)
    return s


def _propagate_expected(o, s, level, jump):
    """calculatePropagateOffset for the four neighbours, test_offset and the
    strict in-order selection, from the level-0-sized base buffer (pixelWrap
    wraps modulo the level-0 size and reads outside the level rectangle)."""
    S = State(o, s, level, frame=0)
    full, _, _ = o.get_output(0, s.w, s.h)
    H, Wd = full.shape
    exw, exh = S.exb.shape[1], S.exb.shape[0]
    sch, scw = S.scale.shape[:2]
    exp = {}
    for y in range(S.lh):
        for x in range(S.lw):
            if S.image_culled(x, y):
                exp[(x, y)] = ("culled", None)
                continue
            ia = float(S.target[y, x, 3])
            cands, ress = [], []
            for tx, ty in ((x + jump, y), (x - jump, y), (x, y + jump), (x, y - jump)):
                n = full[ty % H, tx % Wd]
                nsc = S.scale[min(max(ty, 0), sch - 1), min(max(tx, 0), scw - 1)]
                (b0x, b0y), (b1x, b1y) = S.basis(float(n["theta"]), float(nsc[0]), float(nsc[1]))
                ex = float(n["x"]) + (x - tx) * b0x + (y - ty) * b1x
                ey = float(n["y"]) + (x - tx) * b0y + (y - ty) * b1y
                ex = 0.0 if ex < 0 else (exw - 1.0 if ex >= exw else ex)
                ey = 0.0 if ey < 0 else (exh - 1.0 if ey >= exh else ey)
                c = _cand(ex, ey, S.angle(x, y, ex, ey), 0.0, S.scale[y, x, 0], S.scale[y, x, 1])
                cands.append(c)
                ress.append(S.test_offset(x, y, c, ia))
            win, ok = _select(float(S.res[y, x]), cands, ress)
            exp[(x, y)] = ("search", (S.off[y, x], cands, win, ok))
    return exp


def _check_propagate(o, s, level, jump):
    exp = _propagate_expected(o, s, level, jump)
    o.propagate(level, jump)
    lw, lh = s.w >> level, s.h >> level
    after, _, res_after = o.get_output(level, lw, lh)
    changed = exact = culled = 0
    for (x, y), (kind, data) in exp.items():
        a = after[y, x]
        if kind == "culled":
            culled += 1
            assert (a["x"], a["y"], a["theta"], a["hysteresis"], a["scale_u"], a["scale_v"]) == (0, 0, 0, 0, 1, 1)
            assert res_after[y, x] == 0.0
            continue
        old, cands, win, ok = data
        hits = [k for k in ok if _same_xy(a, old if k == -1 else cands[k])]
        assert hits, (x, y, a, win, ok)
        exact += _same_xy(a, old if win == -1 else cands[win])
        if win >= 0:
            changed += 1
            c = cands[win]
            if _same_xy(a, c):
                assert abs(float(a["theta"]) - float(c["theta"])) < 1e-4, (x, y)
                assert a["scale_u"] == c["scale_u"] and a["scale_v"] == c["scale_v"] and a["hysteresis"] == 0
    return changed, exact, culled, len(exp)


@pytest.mark.parametrize("jump", [1, 4, 16])
def test_propagate_candidates_and_selection_match_float64_level0(jump):
    with OracleContext() as o:
        s = _search_session(o, 0)
        changed, exact, culled, n = _check_propagate(o, s, 0, jump)
    assert changed > n // 4 and exact >= n - 3 and culled == 0


@pytest.mark.parametrize("jump", [3, 8])
def test_propagate_wraps_modulo_level0_size_at_a_coarse_level(jump):
    """Level 1 (12x10 inside a 24x20 buffer): neighbours at x + J >= 12 and
    x - J < 0 (-> 24 + x - J) are read from the level-0 field left outside the
    level rectangle, as pixelWrap does (cudaImageBuffer.h:81-87)."""
    with OracleContext() as o:
        s = _search_session(o, 1)
        full, _, _ = o.get_output(0, 24, 20)
        assert np.any(full["x"][:, 12:] != 0)  # the stale level-0 field is there
        changed, exact, culled, n = _check_propagate(o, s, 1, jump)
    assert changed > n // 5 and exact >= n - 2


def test_propagate_culling_and_exemplar_rejection():
    """Distance transforms in pixel units: pixels above the cull (50) get the
    zero offset with scale (1, 1) and residual 0; candidates whose truncated
    exemplar texel lies above 1.25 * cull are never accepted."""
    with OracleContext() as o:
        s = _search_session(o, 0, dt_cull=True)
        S = State(o, s, 0, frame=0)
        n_cull = sum(S.image_culled(x, y) for y in range(S.lh) for x in range(S.lw))
        assert 0 < n_cull < S.lw * S.lh
        assert (S.exdt > 62.5).mean() > 0.1 and (S.exdt <= 62.5).mean() > 0.1
        changed, exact, culled, n = _check_propagate(o, s, 0, 2)
        assert culled == n_cull and changed > 0 and exact >= n - culled - 3
        after, _, _ = o.get_output(0, 24, 20)
        before = S.off
        moved = (after["x"] != before["x"]) | (after["y"] != before["y"])
        for y, x in zip(*np.nonzero(moved)):
            if S.image_culled(x, y):
                continue
            ex = min(int(after["x"][y, x]), 29)
            ey = min(int(after["y"][y, x]), 27)
            assert S.exdt[ey, ex] <= 62.5, (x, y)


# ---------------------------------------------------------------- scatter
def _clamp_ex(x, y, w, h):
    """clampToExemplar (texSynth_util.h:50-57)."""
    x = 0.0 if x < 0 else (w - 1.0 if x >= w else x)
    y = 0.0 if y < 0 else (h - 1.0 if y >= h else y)
    return x, y


def _scatter_expected(o, s, level):
    """scatter_kernel with the reference's truncated randFloat() (r = theta = 0):
    sample k sits at the running winner + (0.9 * dist_k, 0) with dist in
    float32, dist_0 = max(exemplar w, h), while 0.9 * dist >= 10 and k <
    scatterSamplesCount; then averageOffset; written only below the cache."""
    S = State(o, s, level, frame=0)
    ew, eh = S.exb.shape[1], S.exb.shape[0]
    exp = {}
    for y in range(S.lh):
        for x in range(S.lw):
            if S.image_culled(x, y):
                exp[(x, y)] = ("culled", S.off[y, x])
                continue
            ia = float(S.target[y, x, 3])
            cached = float(S.res[y, x])
            best = float(np.float32(cached + 10e-7))
            cur = S.off[y, x]
            dist = np.float32(max(ew, eh))
            tie = False
            i = 0
            while i < s.params.scatter_samples_count and dist * np.float32(0.9) >= np.float32(10):
                ex, ey = _clamp_ex(float(cur["x"]) + float(dist * np.float32(0.9)), float(cur["y"]), ew, eh)
                c = _cand(ex, ey, S.angle(x, y, ex, ey), 0.0, S.scale[y, x, 0], S.scale[y, x, 1])
                rk = S.test_offset(x, y, c, ia)
                tie |= abs(rk - best) <= 1e-5 * max(best, 1e-9)
                if rk < best:
                    cur, best = c, rk
                dist = dist * np.float32(0.9)
                i += 1
            nb = [(x + i2, y + j2) for i2 in (-1, 0, 1) for j2 in (-1, 0, 1)
                  if (i2, j2) != (0, 0) and 0 <= x + i2 < S.lw and 0 <= y + j2 < S.lh]
            ax = sum(float(S.off[q, p]["x"]) for p, q in nb) / len(nb)
            ay = sum(float(S.off[q, p]["y"]) for p, q in nb) / len(nb)
            avg = _cand(ax, ay, S.angle(x, y, ax, ay), 1.0)
            ra = S.test_offset(x, y, avg, ia)
            tie |= abs(ra - best) <= 1e-5 * max(best, 1e-9)
            if ra < best:
                cur, best = avg, ra
            tie |= abs(best - cached) <= 1e-5 * max(cached, 1e-9)
            out = cur if best < cached else S.off[y, x]
            exp[(x, y)] = ("search", (out, tie, not _same_xy(out, S.off[y, x], 0)))
    return exp


@pytest.mark.parametrize("r,level", [(2, 0), (3, 0), (2, 1)])
def test_scatter_reference_rng_matches_float64(r, level):
    with OracleContext() as o:
        s = _search_session(o, level, r=r)
        exp = _scatter_expected(o, s, level)
        o.scatter(level)
        lw, lh = s.w >> level, s.h >> level
        after, _, _ = o.get_output(level, lw, lh)
    moved = mism = n_avg = 0
    for (x, y), (kind, data) in exp.items():
        a = after[y, x]
        assert kind == "search"
        out, tie, mv = data
        if not _same_xy(a, out):
            assert tie, (x, y, a, out)
            mism += 1
            continue
        moved += mv
        if mv and out["hysteresis"] == 1.0:  # the average candidate won
            n_avg += 1
            assert a["hysteresis"] == 1.0 and a["scale_u"] == 1.0 and a["scale_v"] == 1.0
        elif mv:
            assert a["hysteresis"] == 0.0 and a["scale_u"] == out["scale_u"] and a["scale_v"] == out["scale_v"]
            assert abs(float(a["theta"]) - float(out["theta"])) < 1e-4
    assert mism <= 2 and moved > len(exp) // 5 and n_avg > 0


@pytest.mark.parametrize("ex,expected", [(1024, 25), (512, 25), (256, 25), (128, 24), (64, 17), (30, 10), (11, 0)])
def test_scatter_sample_count_closed_form(ex, expected):
    """n_s = min(samples, #{k : 0.9^(k+1) * W_ex >= 10}) (float32 loop test):
    25, 25, 25, 24, 17 at exemplar 1024..64; read through the nominal
    evaluation count (samples plus the average candidate per pixel)."""
    st = {"base": np.full((ex, ex, 4), 0.5, np.float32), "output": np.full((ex, ex, 4), 0.5, np.float32)}
    with OracleContext() as o:
        s = Session(o, st, [{"color": W.sin_noise(5, 2, 8, 8)}],
                    o.default_params(residual_window_size=1, levels=1, orientation_enabled=0))
        s.run([("style", 0), ("frame", 0, 0), ("randomize",)])
        before = o.pde_count()
        o.scatter(0)
        after = o.pde_count()
    assert after - before == 64 * (expected + 1) + 64  # + the residual refresh of the sync


def test_average_offset_closed_forms_at_borders():
    """averageOffset: the mean of the in-bounds 8-neighbourhood without the
    centre -- on an affine field the pixel's own value inside, shifted by the
    neighbours' mean position at edges (5 neighbours) and corners (3).  The
    cache is set above any energy and no random sample is drawn, so every
    pixel takes its average candidate (hysteresis 1, scale (1, 1))."""
    w, h = 12, 10
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    fx, fy = 3 + 0.5 * xx + 0.25 * yy, 2 + 0.25 * xx + 0.5 * yy
    field = np.zeros((h, w), _ffi.PATCHXF_DTYPE)
    field["x"], field["y"], field["scale_u"], field["scale_v"] = fx, fy, 1, 1
    st = W.style_pair(24, 24, seed=331, guides=True)
    with OracleContext() as o:
        s = Session(o, st, [W.still_frame(w, h, seed=332, guides=True)],
                    o.default_params(residual_window_size=1, levels=1, scatter_samples_count=0, transparency_ok=1))
        s.run([("style", 0), ("frame", 0, 0)])
        o.load_offsets(0, field)
        o.sync_buffers(0, _ffi.PT_DONT_COPY_OFFSETS)
        o.load_residual(0, np.full((h, w), 1e30, np.float32))
        S = State(o, s, 0, frame=0)
        o.scatter(0)
        after, _, _ = o.get_output(0, w, h)
    for y in range(h):
        for x in range(w):
            di = [i for i in (-1, 0, 1) for j in (-1, 0, 1) if (i, j) != (0, 0) and 0 <= x + i < w and 0 <= y + j < h]
            dj = [j for i in (-1, 0, 1) for j in (-1, 0, 1) if (i, j) != (0, 0) and 0 <= x + i < w and 0 <= y + j < h]
            mi, mj = np.mean(di), np.mean(dj)
            ex, ey = fx[y, x] + 0.5 * mi + 0.25 * mj, fy[y, x] + 0.25 * mi + 0.5 * mj
            a = after[y, x]
            assert abs(a["x"] - ex) < 2e-6 * ex and abs(a["y"] - ey) < 2e-6 * ey, (x, y, a, ex, ey)
            assert a["hysteresis"] == 1 and a["scale_u"] == 1 and a["scale_v"] == 1
            assert abs(a["theta"] - S.angle(x, y, float(a["x"]), float(a["y"]))) < 1e-6
    # interior pixels see the field itself; corners are pulled inwards by (2/3, 2/3)
    assert abs(after[5, 5]["x"] - fx[5, 5]) < 1e-6
    assert abs(after[0, 0]["x"] - (fx[0, 0] + 0.5 * 2 / 3 + 0.25 * 2 / 3)) < 1e-5


def test_scatter_skips_culled_pixels():
    """A culled pixel returns before any write (:880-883): its working offset
    stays what the previous step left there."""
    with OracleContext() as o:
        s = _search_session(o, 0, dt_cull=True)
        s.run([("propagate", 2)])
        S = State(o, s, 0, frame=0)
        o.scatter(0)
        after, _, res = o.get_output(0, 24, 20)
    culled = S.imgdt > 50
    assert culled.any()
    assert after[culled].tobytes() == S.off[culled].tobytes()
    assert np.all(after["x"][culled] == 0) and np.all(res[culled] == 0)


# ---------------------------------------------------------------- upsample
def test_upsample_nearest_fallback_with_rotation_and_fine_clamp():
    """Coarse field constant except a few outliers 8 px away: a fine pixel
    whose four coarse corners include one has a corner more than 4 px from the
    bilinear mean, so it takes the rounded coarse offset plus its rotated
    sub-pixel delta, clamped with the FINE exemplar size, times 2 plus 0.5
    (:1403-1417, :1485-1497); the others get 2 * bilinear + 0.5."""
    w, h = 24, 20
    cw, ch = 12, 10
    coarse = np.zeros((ch, cw), _ffi.PATCHXF_DTYPE)
    coarse["x"], coarse["y"], coarse["theta"] = 6.25, 5.5, 0.4
    coarse["scale_u"] = coarse["scale_v"] = 1
    outliers = [(3, 2), (8, 6), (10, 1)]
    for cx, cy in outliers:
        coarse[cy, cx]["x"] = 14.25  # beyond the coarse exemplar (15 wide) minus 1: fine clamp keeps it
        coarse[cy, cx]["theta"] = -0.9
    st = W.style_pair(30, 28, seed=341, guides=True)
    with OracleContext() as o:
        s = Session(o, st, [W.still_frame(w, h, seed=342, guides=True)],
                    o.default_params(residual_window_size=1, levels=2))
        s.run([("style", 0), ("frame", 0, 0)])
        o.load_offsets(1, coarse)
        o.upsample(0)
        S = State(o, s, 0, frame=0)
        off = S.off
    n_near = 0
    for y in range(1, h - 1):
        for x in range(1, w - 1):
            cx, cy = (x - 0.5) / 2, (y - 0.5) / 2
            lx, ly = int(np.floor(cx)), int(np.floor(cy))
            a, b = cx - lx, cy - ly
            q = [coarse[ly, lx], coarse[ly, lx + 1], coarse[ly + 1, lx], coarse[ly + 1, lx + 1]]
            wts = [(1 - a) * (1 - b), a * (1 - b), (1 - a) * b, a * b]
            ix = sum(wk * float(o_["x"]) for wk, o_ in zip(wts, q))
            iy = sum(wk * float(o_["y"]) for wk, o_ in zip(wts, q))
            if any((ix - float(o_["x"])) ** 2 + (iy - float(o_["y"])) ** 2 > 16 for o_ in q):
                n_near += 1
                rx, ry = int(np.floor(cx + 0.5)), int(np.floor(cy + 0.5))  # round half away (cx > 0)
                n = coarse[ry, rx]
                dx, dy = cx - rx, cy - ry
                th = float(n["theta"])
                ux = float(n["x"]) + dx * np.cos(th) - dy * np.sin(th)
                uy = float(n["y"]) + dx * np.sin(th) + dy * np.cos(th)
                ux, uy = min(max(ux, 0), 29), min(max(uy, 0), 27)
                ex, ey = 2 * ux + 0.5, 2 * uy + 0.5
            else:
                ex, ey = 2 * ix + 0.5, 2 * iy + 0.5
            g = off[y, x]
            assert abs(g["x"] - ex) < 1e-5 * ex and abs(g["y"] - ey) < 1e-5 * ey, (x, y, g, ex, ey)
            assert abs(g["theta"] - S.angle(x, y, float(g["x"]), float(g["y"]))) < 1e-6
    assert n_near >= 3 * 9  # every fine pixel around each outlier


# ---------------------------------------------------------------- advection culling
def test_advection_previous_frame_culling_reads_the_truncated_texel():
    """isImagePixelCulledPreviousFrame(tx, ty, level, 8.0) has int
    parameters: the advected source position is truncated and the previous
    frame's distance texel there compared with cull - 8 = 42; a current pixel
    above the cull (50) gets PatchXF() and zero colours."""
    w, h = 24, 20
    vf = np.zeros((h, w, 2), np.float32)
    vf[..., 0], vf[..., 1] = 1.5, 0.5
    dt0 = W.culling_dist_trans(351, 0, w, h)
    dt1 = W.culling_dist_trans(351, 1, w, h, lo=10.0, hi=60.0)
    f0 = {"color": W.sin_noise(352, 2, w, h), "vel_f": vf, "vel_b": np.zeros((h, w, 2), np.float32),
          "dist_trans": dt0}
    f1 = {"color": W.sin_noise(353, 2, w, h), "vel_f": np.zeros((h, w, 2), np.float32), "vel_b": -vf,
          "dist_trans": dt1}
    st = W.style_pair(32, 32, seed=354)
    with OracleContext() as o:
        s = Session(o, st, [f0, f1], o.default_params(residual_window_size=2, levels=1, orientation_enabled=0))
        s.run([("style", 0), ("reset_adv",), ("frame", 0, 0), ("randomize",), ("flip",), ("frame", 0, 1),
               ("advect", 1)])
        adv, base, out, _, ts = o.get_advected(0, 0, w, h)
    n_prev = n_cur = n_ok = 0
    for y in range(h):
        for x in range(w):
            if dt1[y, x] > 50:
                n_cur += 1
                assert adv[y, x]["hysteresis"] == 0 and adv[y, x]["x"] == 0 and np.all(base[y, x] == 0)
                continue
            tx, ty = x - 1.5, y - 0.5
            if not (tx >= 0 and ty >= 0 and np.ceil(tx) < w and np.ceil(ty) < h):
                assert adv[y, x]["hysteresis"] == 0
                continue
            if dt0[int(ty), int(tx)] > 42.0:
                n_prev += 1
                assert adv[y, x]["hysteresis"] == 0 and ts[y, x] == -1, (x, y)
            else:
                n_ok += 1
                assert adv[y, x]["hysteresis"] == 1 and ts[y, x] == 1, (x, y)
    assert n_prev > 10 and n_cur > 10 and n_ok > 10
