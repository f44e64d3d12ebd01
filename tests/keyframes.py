"""Keyframe-interpolation inputs and an operator driver for either backend
(HIP library or CPU oracle), mirroring how synthesisProcessor.cpp drives
TexSynth with keys: the styles are the keyframes (image-sized exemplars,
painted where the output's alpha is above 0), the cumulative key-to-key
fields and the frame-to-key ribbons are RibbonP textures packed as
(time_step, y, x, layer), ribbon pyramids downsample them
(imagePyramid.cpp:211-235)."""
import numpy as np

from painttween_amd import _ffi
from painttween_amd import workloads as W
from tests.harness import first_difference

E1, E2, E3, E4 = _ffi.PT_ELEM_F32, _ffi.PT_ELEM_F32X2, _ffi.PT_ELEM_F32X3, _ffi.PT_ELEM_F32X4
CONST, WITH_LEVELS, RIBBON = (_ffi.PT_DOWNSAMPLE_SCALE_CONSTANT, _ffi.PT_DOWNSAMPLE_SCALE_WITH_LEVELS,
                              _ffi.PT_DOWNSAMPLE_RIBBON)


def key_styles(w, h, nk, seed):
    """nk keyframe styles: input/output pairs with guides; each output is
    painted (alpha in [0.5, 1]) on smooth blobs covering about half of the
    frame and transparent (alpha 0) elsewhere."""
    styles = []
    for k in range(nk):
        st = W.style_pair(w, h, seed * 7 + k, guides=True)
        mask = W.smooth_noise(seed, 300 + k, w, h) > 0.5
        out = st["output"].copy()
        out[..., 3] = np.where(mask, 0.5 + 0.5 * out[..., 3], 0.0)
        st["output"] = out
        styles.append(st)
    return styles


def ribbon_texels(w, h, nk, seed, stream, invalid=0.15, max_d=4.0, ts_lo=-1, ts_hi=4):
    """A RibbonP texture (time_step, y, x, layer): smooth displacements of up
    to max_d pixels, integer time steps in [ts_lo, ts_hi) (a negative step
    stops a propagation walk), layers in [0, nk) and about `invalid` of the
    texels invalid (layer -1, which stops a scatter walk)."""
    rng = np.random.default_rng([int(seed), int(stream), 0xB1B])
    t = np.empty((h, w, 4), np.float32)
    t[..., 0] = rng.integers(ts_lo, ts_hi, (h, w))
    t[..., 1] = max_d * (2.0 * W.smooth_noise(seed, 400 + stream, w, h) - 1.0)
    t[..., 2] = max_d * (2.0 * W.smooth_noise(seed, 500 + stream, w, h) - 1.0)
    lay = rng.integers(0, nk, (h, w))
    t[..., 3] = np.where(rng.uniform(0.0, 1.0, (h, w)) < invalid, -1, lay)
    return t


def ribbon_field(w, h, nk, seed, stream, invalid=0.3):
    """A level of the ribbon field as pt_ribbonp records."""
    t = ribbon_texels(w, h, nk, seed, stream, invalid=invalid, ts_lo=-3, ts_hi=5)
    r = np.zeros((h, w), _ffi.RIBBON_DTYPE)
    r["x"], r["y"], r["time_step"], r["layer"] = t[..., 2], t[..., 1], t[..., 0].astype(np.int32), \
        t[..., 3].astype(np.int32)
    return r


def key_params(ctx, nk, **overrides):
    kw = dict(residual_window_size=2, levels=2, num_styles=nk, interpolate_key_frame=1, first_key_index=0,
              last_key_index=nk - 1, transparency_ok=1, orientation_enabled=1, dist_trans_weight=0.5,
              rng_mode=_ffi.PT_RNG_COUNTER, seed=77, scatter_samples_count=9)
    kw.update(overrides)
    return ctx.default_params(**kw)


class KeySession:
    """Pyramids of nk keyframe styles, frames and their ribbons on one
    backend context."""

    def __init__(self, ctx, styles, frames, params, seed=1):
        self.ctx, self.params = ctx, params
        nk = len(styles)
        h, w = frames[0]["color"].shape[:2]
        sh, sw = styles[0]["base"].shape[:2]
        self.w, self.h, self.sw, self.sh, self.nk = w, h, sw, sh, nk
        ctx.setup(params)
        ctx.initialize(w, h, sw, sh)
        self.styles = []
        for k, st in enumerate(styles):
            t = ctx.pyramid(E3, st["orientation"])
            self.styles.append({
                "base": ctx.pyramid(E4, st["base"]), "output": ctx.pyramid(E4, st["output"]),
                "tensor": t, "orient": ctx.orientation_pyramid(t), "dt": ctx.pyramid(E1, st["dist_trans"]),
                # key-to-key fields: layer k of B leads from key k to key k-1, of F to key k+1
                "cumB": ctx.pyramid(E4, ribbon_texels(sw, sh, nk, seed, 10 + k), RIBBON),
                "cumF": ctx.pyramid(E4, ribbon_texels(sw, sh, nk, seed, 20 + k), RIBBON),
            })
        self.frames = []
        for i, f in enumerate(frames):
            tp = ctx.pyramid(E3, f["orientation"])
            self.frames.append({
                "color": ctx.pyramid(E4, f["color"]), "tensor": tp, "orient": ctx.orientation_pyramid(tp),
                "dt": ctx.pyramid(E1, f["dist_trans"], default=1.0),
                "scale": ctx.pyramid(E2, None, shape=(w, h), default=1.0),
                "velb": ctx.pyramid(E2, f["vel_b"], WITH_LEVELS), "velf": ctx.pyramid(E2, f["vel_f"], WITH_LEVELS),
                "ftkB": ctx.pyramid(E4, ribbon_texels(w, h, nk, seed, 30 + 2 * i), RIBBON),
                "ftkF": ctx.pyramid(E4, ribbon_texels(w, h, nk, seed, 31 + 2 * i), RIBBON),
            })
        self.level = None

    def load_styles(self, level):
        c = self.ctx
        lw, lh = self.sw >> level, self.sh >> level
        c.initialize_exemplars(lw, lh, self.nk)
        c.initialize_cumulative_fields(lw, lh, self.nk)
        for k, s in enumerate(self.styles):
            c.upload_image(level, "EXEMPLAR_BASE", s["base"], k)
            c.upload_image(level, "EXEMPLAR_OUTPUT", s["output"], k)
            c.upload_image(level, "EXEMPLAR_ORIENTATION", s["orient"], k)
            c.upload_image(level, "EXEMPLAR_DIST_TRANS", s["dt"], k)
            c.upload_image(level, "CUMUL_FIELD_B", s["cumB"], k)
            c.upload_image(level, "CUMUL_FIELD_F", s["cumF"], k)
        self.level = level

    def load_frame(self, level, f, random=True, ribbons=True):
        c = self.ctx
        fp = self.frames[f]
        c.upload_image(level, "INPUT_COLOR", fp["color"])
        c.upload_image(level, "INPUT_ORIENTATION", fp["orient"])
        c.upload_image(level, "INPUT_VEL_B", fp["velb"])
        c.upload_image(level, "INPUT_VEL_F", fp["velf"])
        c.upload_image(level, "INPUT_DIST_TRANS", fp["dt"])
        c.upload_image(level, "INPUT_SCALE", fp["scale"])
        if ribbons:
            c.upload_image(level, "FRAME_TO_KEY_RIBBON_B", fp["ftkB"])
            c.upload_image(level, "FRAME_TO_KEY_RIBBON_F", fp["ftkF"])
        if random:
            c.create_random_pyramid_keys(level, f, fp["color"], fp["orient"],
                                         [s["output"] for s in self.styles], [s["orient"] for s in self.styles])
        self.level = level

    def state(self):
        """Base offsets (layer included), canvas, residual, both advected
        sets and the ribbon field of the current level."""
        c, l = self.ctx, self.level
        lw, lh = self.w >> l, self.h >> l
        out = list(c.get_output(l, lw, lh))
        for k in (0, 1):
            out += list(c.get_advected(l, k, lw, lh))
        out.append(c.get_ribbon_field(l, lw, lh))
        return out

    def run(self, ops):
        out = []
        c = self.ctx
        for op in ops:
            name, args = op[0], op[1:]
            if name == "styles":
                self.load_styles(*args)
            elif name == "frame":
                self.load_frame(*args)
            elif name == "params":
                for k, v in args[0].items():
                    setattr(self.params, k, v)
                c.upload_parameters(self.params)
            elif name == "flip":
                c.flip_anim_buffers()
            elif name == "reset_adv":
                c.reset_advected_buffers()
            elif name == "randomize":
                c.randomize_offsets(self.level)
            elif name == "propagate":
                c.propagate(self.level, args[0], args[1] if len(args) > 1 else -1)
            elif name == "scatter":
                c.scatter(self.level, args[0] if args else -1)
            elif name == "upsample":
                c.upsample(self.level)
            elif name == "sync":
                c.sync_buffers(self.level, args[0] if args else 0)
            elif name == "residual":
                c.update_residual_cache(self.level)
            elif name == "advect":
                c.advect_offsets(self.level, args[1], args[0], args[2] if len(args) > 2 else 0)
            elif name == "merge_random":
                c.merge_randomly(self.level, args[0], args[1] if len(args) > 1 else 0.5)
            elif name == "merge_residual":
                c.merge_using_residual(self.level, args[0])
            elif name == "merge_advected":
                c.merge_advected_buffers(self.level, args[0])
            elif name == "ftk_offsets":
                c.offsets_from_frame_to_key_ribbon(self.level, args[0])
            elif name == "load_ribbons":
                lw, lh = self.w >> self.level, self.h >> self.level
                c.load_ribbon_field(self.level, ribbon_field(lw, lh, self.nk, *args))
            elif name == "accumulate":
                c.accumulate_ribbon_field(self.level, args[0], args[1], args[2])
            elif name == "load_offsets":
                c.load_offsets(self.level, args[0])
            else:
                raise ValueError(name)
            out.append((name, self.state()))
        return out


LABELS = ("offsets", "canvas", "residual", "advF.offsets", "advF.base", "advF.output", "advF.residual",
          "advF.time_step", "advB.offsets", "advB.base", "advB.output", "advB.residual", "advB.time_step",
          "ribbons")


def assert_same_key_states(ga, oa):
    assert len(ga) == len(oa)
    for i, ((n1, s1), (n2, s2)) in enumerate(zip(ga, oa)):
        for label, x, y in zip(LABELS, s1, s2):
            d = first_difference(x, y)
            assert d is None, "step %d (%s): %s %s" % (i, n1, label, d)
