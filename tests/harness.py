"""Drive any backend (HIP library or CPU oracle) through the same TexSynth
operator sequence and record the state after every step."""
import numpy as np

from painttween_amd import _ffi

E1, E2, E3, E4 = _ffi.PT_ELEM_F32, _ffi.PT_ELEM_F32X2, _ffi.PT_ELEM_F32X3, _ffi.PT_ELEM_F32X4
I32 = _ffi.PT_ELEM_I32
CONST, WITH_LEVELS = _ffi.PT_DOWNSAMPLE_SCALE_CONSTANT, _ffi.PT_DOWNSAMPLE_SCALE_WITH_LEVELS


class Session:
    """Pyramids for one style and a list of frames on one backend context."""

    def __init__(self, ctx, style, frames, params):
        self.ctx = ctx
        self.params = params
        sh, sw = style["base"].shape[:2]
        h, w = frames[0]["color"].shape[:2]
        self.w, self.h, self.sw, self.sh = w, h, sw, sh
        ctx.setup(params)
        ctx.initialize(w, h, sw, sh)
        self.sb = ctx.pyramid(E4, style["base"])
        self.so = ctx.pyramid(E4, style["output"])
        self.st = ctx.pyramid(E3, style.get("orientation"), shape=(sw, sh))
        self.sor = ctx.orientation_pyramid(self.st)
        self.sdt = ctx.pyramid(E1, style.get("dist_trans"), shape=(sw, sh), default=0.0)
        self.sid = None if style.get("object_ids") is None else ctx.pyramid(I32, style["object_ids"])
        self.frames = []
        for f in frames:
            fp = {
                "color": ctx.pyramid(E4, f["color"]),
                "tensor": ctx.pyramid(E3, f.get("orientation"), shape=(w, h)),
                "dt": ctx.pyramid(E1, f.get("dist_trans"), shape=(w, h), default=1.0),
                "scale": ctx.pyramid(E2, f.get("scale"), shape=(w, h), default=1.0),
                "velb": ctx.pyramid(E2, f.get("vel_b"), WITH_LEVELS, shape=(w, h)),
                "velf": ctx.pyramid(E2, f.get("vel_f"), WITH_LEVELS, shape=(w, h)),
                "id": None if f.get("object_id") is None else ctx.pyramid(I32, f["object_id"]),
            }
            fp["orient"] = ctx.orientation_pyramid(fp["tensor"])
            self.frames.append(fp)
        self.level = None

    def load_style(self, level):
        c = self.ctx
        c.initialize_exemplars(self.sw >> level, self.sh >> level, 1)
        c.upload_image(level, "EXEMPLAR_BASE", self.sb)
        c.upload_image(level, "EXEMPLAR_OUTPUT", self.so)
        c.upload_image(level, "EXEMPLAR_ORIENTATION", self.sor)
        c.upload_image(level, "EXEMPLAR_DIST_TRANS", self.sdt)
        if self.sid is not None:
            c.upload_image(level, "EXEMPLAR_OBJECT_IDS", self.sid)
        self.level = level

    def load_frame(self, level, f, random=True, frame_number=None):
        c = self.ctx
        fp = self.frames[f]
        c.upload_image(level, "INPUT_COLOR", fp["color"])
        c.upload_image(level, "INPUT_ORIENTATION", fp["orient"])
        if fp["id"] is not None:
            c.upload_image(level, "INPUT_ID", fp["id"])
        c.upload_image(level, "INPUT_VEL_B", fp["velb"])
        c.upload_image(level, "INPUT_VEL_F", fp["velf"])
        c.upload_image(level, "INPUT_DIST_TRANS", fp["dt"])
        c.upload_image(level, "INPUT_SCALE", fp["scale"])
        if random:
            c.create_random_pyramid(level, f if frame_number is None else frame_number, fp["color"],
                                    fp["orient"], self.so, self.sor)
        self.level = level

    def state(self, level=None):
        level = self.level if level is None else level
        return self.ctx.get_output(level, self.w >> level, self.h >> level)

    def run(self, ops):
        """Apply ops; return the list of states (offsets, canvas, residual)
        after each op."""
        out = []
        c = self.ctx
        for op in ops:
            name, args = op[0], op[1:]
            if name == "style":
                self.load_style(*args)
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
                c.propagate(self.level, args[0])
            elif name == "scatter":
                c.scatter(self.level)
            elif name == "upsample":
                c.upsample(self.level)
            elif name == "sync":
                c.sync_buffers(self.level, args[0] if args else 0)
            elif name == "residual":
                c.update_residual_cache(self.level)
            elif name == "advect":
                c.advect_offsets(self.level, 1, args[0], args[1] if len(args) > 1 else 0)
            elif name == "merge_random":
                c.merge_randomly(self.level, args[0], args[1] if len(args) > 1 else 0.5)
            elif name == "merge_residual":
                c.merge_using_residual(self.level, args[0])
            else:
                raise ValueError(name)
            out.append((name, self.state()))
        return out


def first_difference(a, b):
    """Index and values of the first differing element of two arrays, or None."""
    ab = np.frombuffer(np.ascontiguousarray(a).tobytes(), np.uint32)
    bb = np.frombuffer(np.ascontiguousarray(b).tobytes(), np.uint32)
    if ab.shape != bb.shape:
        return "shape %s vs %s" % (a.shape, b.shape)
    d = np.nonzero(ab != bb)[0]
    if len(d) == 0:
        return None
    return "%d words differ; first at word %d: %s vs %s" % (
        len(d), d[0], np.frombuffer(ab[d[0]:d[0] + 1].tobytes(), np.float32),
        np.frombuffer(bb[d[0]:d[0] + 1].tobytes(), np.float32))


def assert_same_states(ga, oa):
    for i, ((n1, s1), (n2, s2)) in enumerate(zip(ga, oa)):
        for label, x, y in zip(("offsets", "canvas", "residual"), s1, s2):
            diff = first_difference(x, y)
            assert diff is None, "step %d (%s): %s %s" % (i, n1, label, diff)
