"""Working sets: the reference's XML description of a shot, its styles and the
synthesis parameters (NLSynthesizer::readWorkingSet / writeWorkingSet,
src/PaintTweenGUI/nlSynthesizer.cpp:101-194), read into what the HIP path
takes: a `pt_params`, the schedule settings and the frame and style images.

Layout of a working set (a `doc` root element):

    <styles><style [frame="k"]><styleInput path=".."/><styleOutput path=".."/>
            [<styleOrientation/>] [<styleDistanceTransform/>] [<styleObjectIDs/>]</style>..</styles>
    <params><coherenceWeight value=".."/> ... <offsetsHistogramSlope value0=".." ../></params>
    <anim><unit name=".."/><shot name=".."/><shotDir path=".."/><workingDir path=".."/>
          <elements inputElem=".." orientationElem=".." .. velBElem=".." velFElem=".."/>
          <fileAccess checkFarm="0|1" checkFiles=".."/></anim>

Element tables follow nlInputElements.cpp:30-42 and nlStyles.cpp:31-38,
parameter names and defaults nlParameters.cpp:34-138, the shot rules
nlShot.cpp:55-194, and the translation into TsParameters
SynthesisProcessor::updateParameters (src/PaintTweenCUDA/synthesisProcessor.cpp:167-262).
Images are OpenEXR files read with painttween_amd.exr: the colour input as
an RGBA image rounded to half floats (RgbaInputFile, imageIO.cpp:87-130),
every other element as FP32 channels named `<layer>.<channel>`, a channel
the file lacks reading as 0 (imageIO.cpp:213-260).
"""
import os
import re
import tempfile
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

import numpy as np

from . import _ffi, exr

# (key, file layer, XML attribute, channels) -- nlInputElements.cpp:30-42
INPUT_ELEMENTS = (
    ("input", "rgba", "inputElem", ("red", "green", "blue", "alpha")),
    ("orientation", "orientation", "orientationElem", ("Ix2", "IxIy", "Iy2")),
    ("dist_trans", "dist_trans", "distTransElem", ("length",)),
    ("scale", "scale", "scaleElem", ("x", "y")),
    ("id", "id", "idMergedElem", ("red", "green", "blue")),
    ("surf_id", "surf_id", "surfIdElem", ("u", "v")),
    ("vel_b", "vel_b_1", "velBElem", ("x", "y")),
    ("vel_f", "vel_f_1", "velFElem", ("x", "y")),
    ("ribbon_b", "vel_b_%1", "ribbonBElem", ("x", "y")),
    ("ribbon_f", "vel_f_%1", "ribbonFElem", ("x", "y")),
)
_RIBBON_START = 8  # elements from here on are ribbons, not per-frame inputs

# (key, file layer, XML element, channels) -- nlStyles.cpp:31-38
STYLE_ELEMENTS = (
    ("output", "style_output", "styleOutput", ("red", "green", "blue", "alpha")),
    ("input", "style_input", "styleInput", ("red", "green", "blue", "alpha")),
    ("dist_trans", "dist_trans", "styleDistanceTransform", ("length",)),
    ("orientation", "orientation", "styleOrientation", ("Ix2", "IxIy", "Iy2")),
    ("object_ids", "id", "styleObjectIDs", ("red", "green", "blue")),
)

# names and defaults -- nlParameters.cpp:34-45,105-138 (getTsDefaultParams :63-103)
FLOAT_NAMES = ("coherenceWeight", "hysteresisWeight", "maxDistortion", "inputAnalogyWeight",
               "canvasOutputWeight", "distanceTransformWeight", "edgeBlendingWeight",
               "timeDerivativeInputWeight", "timeDerivativeOutputWeight", "temporalCoherenceWeight",
               "styleObjectIdWeight", "offsetsHistogramSlope", "offsetsHistogramThreshold",
               "advectionJitter", "lumShiftWeight", "alphaChannelWeight", "residualWindowScaleU",
               "residualWindowScaleV")
BOOL_NAMES = ("orientationEnabled", "keyFrameInterpolation", "guidedInterpolation", "lumShiftEnabled",
              "transparencyAllowed")
INT_NAMES = ("version", "iterations", "iterationsFirstFrame", "firstPassDirection", "levels", "maxLevel",
             "scatterSamplesCount", "fixupPasses", "residualWindowSize", "firstFrame", "lastFrame",
             "colorspace", "synthesisScheme")
NUM_MAX_STYLES = 5
DEFAULTS = {
    "advectionJitter": 0.0, "alphaChannelWeight": 1.0, "canvasOutputWeight": 1.0, "coherenceWeight": 0.0,
    "colorspace": 0, "distanceTransformWeight": 0.0, "inputAnalogyWeight": 1.0,
    "keyFrameInterpolation": False, "levels": 4, "lumShiftEnabled": False, "lumShiftWeight": 0.0,
    "orientationEnabled": True, "residualWindowScaleU": 1.0, "residualWindowScaleV": 1.0,
    "residualWindowSize": 4, "scatterSamplesCount": 25, "timeDerivativeInputWeight": 1.0,
    "timeDerivativeOutputWeight": 1.0, "temporalCoherenceWeight": 0.2, "hysteresisWeight": 0.0,
    "maxDistortion": 4.0, "transparencyAllowed": False, "styleObjectIdWeight": 0.0,
    "guidedInterpolation": False, "firstPassDirection": 0, "synthesisScheme": _ffi.PT_S_FMBM_RM,
    "fixupPasses": 1, "maxLevel": 0, "iterations": 3, "iterationsFirstFrame": 3,
}


class WorkingSetError(ValueError):
    pass


def _qt_float(s):
    """QString::toFloat: 0 when the text is not a number."""
    try:
        return float(s)
    except (TypeError, ValueError):
        return 0.0


def _qt_int(s):
    """QString::toInt: 0 when the text is not a base-10 integer."""
    try:
        return int(s)
    except (TypeError, ValueError):
        return 0


def remap_file_path(path):
    """ImageIO::remapFilePath (imageIO.cpp:52-69): /scratch/tmp/ -> the temp directory."""
    return path.replace("/scratch/tmp/", tempfile.gettempdir() + "/")


def _add_dots(name):
    return "." + name + "." if name else name


def _rm_dots(name):
    if name.startswith("."):
        name = name[1:]
    if name.endswith("."):
        name = name[:-1]
    return name


def _child(elt, tag, what):
    c = elt.find(tag)
    if c is None:
        raise WorkingSetError("%s tag not found in the working set" % what)
    return c


@dataclass
class WorkingSet:
    path: str
    creator: str = ""
    render_crews: str = ""
    params: dict = field(default_factory=lambda: dict(DEFAULTS))
    histogram_slopes: list = field(default_factory=lambda: [0.0] * NUM_MAX_STYLES)
    histogram_thresholds: list = field(default_factory=lambda: [0.0] * NUM_MAX_STYLES)
    styles: list = field(default_factory=list)      # per style: {style key: absolute path}
    key_frames: list = field(default_factory=list)  # frame of each keyed style
    unit: str = ""
    shot: str = ""
    shot_dir: str = ""
    working_dir: str = ""
    elements: dict = field(default_factory=dict)    # input key -> element name
    from_farm: bool = False
    input_paths: dict = field(default_factory=dict)  # input key -> path with "%1" for the frame
    ribbons_b: list = field(default_factory=list)
    ribbons_f: list = field(default_factory=list)

    # -- parameters ------------------------------------------------------
    def get(self, name):
        """NLParameters::get{Float,Int,Bool}: an unknown name reads as 0."""
        return self.params.get(name, 0)

    @property
    def first_frame(self):
        return int(self.get("firstFrame"))

    @property
    def last_frame(self):
        return int(self.get("lastFrame"))

    def pt_params(self, base=None):
        """TsParameters as SynthesisProcessor::updateParameters fills them
        (synthesisProcessor.cpp:176-223), written into `base` (a PtParams,
        default: the library's pt_params_default) and returned."""
        from .texsynth import default_params
        g = self.get
        p = default_params() if base is None else base
        p.residual_window_size = int(g("residualWindowSize"))
        p.coherence_weight = g("coherenceWeight")
        p.dist_trans_weight = g("distanceTransformWeight")
        p.input_analogy_weight = g("inputAnalogyWeight")
        p.canvas_output_weight = g("canvasOutputWeight")
        p.scatter_samples_count = int(g("scatterSamplesCount"))
        p.time_derivative_input_weight = g("timeDerivativeInputWeight")
        p.time_derivative_output_weight = g("timeDerivativeOutputWeight")
        p.temporal_coherence_weight = g("temporalCoherenceWeight")
        p.hysteresis_weight = g("hysteresisWeight")
        p.max_distortion = g("maxDistortion")
        p.advection_jitter = float(np.float32(g("advectionJitter"))) * 0.01  # percent; double product
        n = len(self.styles)
        if n >= NUM_MAX_STYLES:
            raise WorkingSetError("%d styles: at most %d" % (n, NUM_MAX_STYLES - 1))
        for i in range(n):
            p.offsets_histogram_slope[i] = self.histogram_slopes[i]
            p.offsets_histogram_threshold[i] = self.histogram_thresholds[i]
        p.num_styles = n
        p.alpha_weight = g("alphaChannelWeight")
        p.transparency_ok = int(bool(g("transparencyAllowed")))
        p.colorspace = int(g("colorspace"))
        p.orientation_enabled = int(bool(g("orientationEnabled")))
        p.style_obj_id_weight = g("styleObjectIdWeight")
        p.lum_shift_enable = int(bool(g("lumShiftEnabled")))
        p.lum_shift_weight = g("lumShiftWeight")
        p.levels = int(g("levels"))
        p.residual_window_scale_u = g("residualWindowScaleU")
        p.residual_window_scale_v = g("residualWindowScaleV")
        p.dist_trans_cull = 50
        p.num_frames = self.last_frame - self.first_frame
        p.use_guide = int(bool(g("guidedInterpolation")))
        p.interpolate_key_frame = int(bool(g("keyFrameInterpolation")))
        p.first_pass_direction = int(g("firstPassDirection"))  # synthesisProcessor.cpp:262
        p.max_level = int(g("maxLevel"))  # synthesisProcessor.cpp:227 (_finalLevel)
        return p

    def schedule(self):
        """The processor settings of an offline render (synthesisProcessor.cpp:225-262):
        scheme, iterations per level, fixup passes and the frame range (the
        first pass direction and maxLevel travel in pt_params).  Settings this
        path does not implement are rejected here; parameters it does not
        implement are rejected by pt_setup or pt_synthesize."""
        g = self.get
        if g("keyFrameInterpolation"):
            raise WorkingSetError("keyFrameInterpolation is not implemented")
        return dict(scheme=int(g("synthesisScheme")), iterations=int(g("iterations")),
                    fixup_passes=int(g("fixupPasses")), first_frame=self.first_frame,
                    last_frame=self.last_frame)

    # -- images ----------------------------------------------------------
    def input_path(self, key, frame):
        return self.input_paths[key].replace("%1", str(frame))

    def read_input(self, key, frame):
        """One frame's input element as float32 (H, W, C), or None when the
        shot does not have it."""
        if key not in self.input_paths:
            return None
        desc = next(d for d in INPUT_ELEMENTS if d[0] == key)
        return _read_layer(self.input_path(key, frame), desc[1], desc[3])

    def read_frame(self, frame):
        """The guides of `frame` in the form Context.synthesize takes."""
        get = self.read_input
        dt = get("dist_trans", frame)
        ids = get("id", frame)
        return {"color": get("input", frame), "orientation": get("orientation", frame),
                "dist_trans": None if dt is None else dt[..., 0], "scale": get("scale", frame),
                "vel_b": get("vel_b", frame), "vel_f": get("vel_f", frame),
                "object_id": None if ids is None else ids_from_rgb(ids, style=False)}

    def read_style(self, index=0):
        """Style `index` in the form Context.synthesize takes."""
        st = self.styles[index]
        out = {}
        for key, layer, _, chans in STYLE_ELEMENTS:
            if key in st:
                out[key] = _read_layer(st[key], layer, chans)
        dt = out.get("dist_trans")
        ids = out.get("object_ids")
        return {"base": out["input"], "output": out["output"], "orientation": out.get("orientation"),
                "dist_trans": None if dt is None else dt[..., 0],
                "object_ids": None if ids is None else ids_from_rgb(ids, style=True)}

    # -- writing ---------------------------------------------------------
    def save(self, path):
        """NLSynthesizer::writeWorkingSet: anim, styles, params (nlShot.cpp:196-228,
        nlStyles.cpp:84-105, nlParameters.cpp:194-228).  Paths are written as held
        (absolute)."""
        root = ET.Element("doc", creator=self.creator)
        anim = ET.SubElement(root, "anim")
        ET.SubElement(anim, "unit", name=self.unit)
        ET.SubElement(anim, "shot", name=self.shot)
        ET.SubElement(anim, "shotDir", path=self.shot_dir)
        ET.SubElement(anim, "workingDir", path=self.working_dir)
        el = ET.SubElement(anim, "elements")
        for key, _, xml_name, _ in INPUT_ELEMENTS:
            if key in self.elements:
                el.set(xml_name, self.elements[key])
        ET.SubElement(anim, "fileAccess", checkFarm="1" if self.from_farm else "0",
                      checkFiles="0" if self.from_farm else "1")
        styles = ET.SubElement(root, "styles")
        for i, st in enumerate(self.styles):
            s = ET.SubElement(styles, "style")
            if i < len(self.key_frames):
                s.set("frame", str(self.key_frames[i]))
            for key, _, xml_name, _ in STYLE_ELEMENTS:
                if key in st:
                    ET.SubElement(s, xml_name, path=st[key])
        params = ET.SubElement(root, "params")
        for name in FLOAT_NAMES:
            if name == "offsetsHistogramSlope":
                ET.SubElement(params, name, {"value%d" % i: repr(v) for i, v in enumerate(self.histogram_slopes)})
            elif name == "offsetsHistogramThreshold":
                ET.SubElement(params, name,
                              {"value%d" % i: repr(v) for i, v in enumerate(self.histogram_thresholds)})
            elif name in self.params:
                ET.SubElement(params, name, value=repr(float(self.params[name])))
        for name in BOOL_NAMES:
            if name in self.params:
                ET.SubElement(params, name, value=str(int(bool(self.params[name]))))
        for name in INT_NAMES:
            if name in self.params:
                ET.SubElement(params, name, value=str(int(self.params[name])))
        ET.ElementTree(root).write(path, encoding="utf-8", xml_declaration=False)


def ids_from_rgb(rgb, style):
    """Object IDs from an RGB ID image: r + 255 g + 255^2 b of the channels
    scaled to 0..255, rounded half up for shot inputs (floorf(v*255 + 0.5),
    imageIO.cpp:298-317) and half away from zero for styles (round(v*255),
    G/nlDataAccess.cpp:378-395)."""
    v = np.asarray(rgb, np.float32)[..., :3] * np.float32(255)  # the float product
    if style:
        # round() of the float: half away from zero, exact in float64 (a
        # float32 v + 0.5 would round 0.49999997 up to 1)
        d = v.astype(np.float64)
        q = np.sign(d) * np.floor(np.abs(d) + 0.5)
    else:
        q = np.floor(v + np.float32(0.5))  # floorf(v * 255 + 0.5f) in float
    q = q.astype(np.int64)
    return (q[..., 0] + q[..., 1] * 255 + q[..., 2] * 255 * 255).astype(np.int32)


def _read_layer(path, layer, channels):
    planes, hdr = exr.read_exr(path)
    h, w = hdr["height"], hdr["width"]
    if layer == "rgba":
        # RgbaInputFile: R, G, B, A as half; a missing A reads as 1, a missing colour as 0
        names = ("R", "G", "B", "A")
        fill = (0.0, 0.0, 0.0, 1.0)
        out = np.empty((h, w, 4), np.float32)
        for k, (n, f) in enumerate(zip(names, fill)):
            out[..., k] = planes[n].astype(np.float16) if n in planes else f
        return out
    out = np.zeros((h, w, len(channels)), np.float32)
    for k, c in enumerate(channels):
        key = layer + "." + c
        if key in planes:
            out[..., k] = planes[key]
    return out


def load(path, check_files=True):
    """NLSynthesizer::readWorkingSet.  With check_files, an input element with
    a frame missing in [firstFrame, lastFrame] is dropped (the colour input
    is required) and ribbon steps are listed while every frame exists, as
    NLShot::setup does (nlShot.cpp:128-190)."""
    path = os.path.abspath(path)
    try:
        doc = ET.parse(path).getroot()
    except ET.ParseError as e:
        raise WorkingSetError("%s: %s" % (path, e)) from None
    ws = WorkingSet(path=path, creator=doc.get("creator", ""), render_crews=doc.get("render_crews", ""))
    root_dir = os.path.dirname(path)

    def resolve(p):
        return os.path.normpath(os.path.join(root_dir, p))

    # styles (nlStyles.cpp:47-82)
    styles = doc.find("styles")
    if styles is None:
        raise WorkingSetError("no style data found")
    for st in styles:
        if "frame" in st.attrib:
            ws.key_frames.append(_qt_int(st.get("frame")))
        images = {}
        for key, _, xml_name, _ in STYLE_ELEMENTS:
            c = st.find(xml_name)
            if c is not None and "path" in c.attrib:
                images[key] = resolve(c.get("path"))
        if images:
            ws.styles.append(images)
        if not ws.styles or "input" not in ws.styles[-1]:
            raise WorkingSetError("no input style found")
        if "output" not in ws.styles[-1]:
            raise WorkingSetError("no output style found")
    if not ws.styles:
        raise WorkingSetError("no style data found")

    # params (nlParameters.cpp:146-191): defaults, then what the file names
    pe = doc.find("params")
    if pe is not None:
        n = len(ws.styles)
        for name in FLOAT_NAMES:
            c = pe.find(name)
            if c is None:
                continue
            if name in ("offsetsHistogramSlope", "offsetsHistogramThreshold"):
                vals = ws.histogram_slopes if name == "offsetsHistogramSlope" else ws.histogram_thresholds
                del vals[n:]
                vals.extend([0.0] * (n - len(vals)))
                for i in range(n):
                    if "value%d" % i in c.attrib:
                        vals[i] = _qt_float(c.get("value%d" % i))
            else:
                ws.params[name] = _qt_float(c.get("value"))
        for name in BOOL_NAMES:
            c = pe.find(name)
            if c is not None:
                ws.params[name] = _qt_int(c.get("value")) == 1
        for name in INT_NAMES:
            c = pe.find(name)
            if c is not None:
                ws.params[name] = _qt_int(c.get("value"))

    # shot (nlShot.cpp:55-194)
    anim = doc.find("anim")
    if anim is None:
        raise WorkingSetError("no animation data found")
    for tag, attr, dest in (("unit", "name", "unit"), ("shot", "name", "shot"), ("shotDir", "path", "shot_dir"),
                            ("workingDir", "path", "working_dir")):
        c = _child(anim, tag, tag)
        if attr not in c.attrib:
            raise WorkingSetError("%s %s not found in the working set" % (tag, attr))
        v = c.get(attr)
        setattr(ws, dest, resolve(v) if attr == "path" else v)
    el = _child(anim, "elements", "elements")
    for key, _, xml_name, _ in INPUT_ELEMENTS:
        if xml_name in el.attrib:
            ws.elements[key] = _rm_dots(el.get(xml_name))
    fa = _child(anim, "fileAccess", "fileAccess")
    ws.from_farm = _qt_int(fa.get("checkFarm", "0")) == 1

    root = ws.shot_dir if ws.from_farm else os.path.normpath(ws.shot_dir + "/" + ws.shot)
    for key, _, _, _ in INPUT_ELEMENTS[:_RIBBON_START]:
        name = ws.elements.get(key, "")
        if not name:
            continue
        p = root
        if ws.from_farm:
            p = p + "/" + name + "/" + ws.shot
        p = remap_file_path(p + _add_dots(name) + "%1.exr")
        if check_files:
            missing = [f for f in range(ws.first_frame, ws.last_frame + 1)
                       if not os.path.exists(p.replace("%1", str(f)))]
            if missing:
                if key == "input":
                    raise WorkingSetError("cannot load input %s" % p.replace("%1", str(missing[0])))
                continue
        ws.input_paths[key] = p
    for i in range(10):
        step = "%02d" % (1 << i)
        eb = re.sub(r"\[.*\]", step, ws.elements.get("ribbon_b", ""))
        ef = re.sub(r"\[.*\]", step, ws.elements.get("ribbon_f", ""))
        if not eb or not ef:
            continue
        if ws.from_farm:
            eb = root + "/" + eb + "/" + ws.shot + "." + eb
            ef = root + "/" + ef + "/" + ws.shot + "." + ef
        eb += ".%1.exr"
        ef += ".%1.exr"
        if check_files and any(not (os.path.exists(eb.replace("%1", str(f))) and
                                    os.path.exists(ef.replace("%1", str(f))))
                               for f in range(ws.first_frame, ws.last_frame + 1)):
            break
        ws.ribbons_f.append(ef)
        ws.ribbons_b.append(eb)
    return ws


def synthesize(ctx, ws, want=("offsets", "canvas", "residual")):
    """Render the working set's frame range with its first style on `ctx`
    (a painttween_amd.Context): the offline render of SynthesisProcessor.
    Returns what Context.synthesize returns."""
    s = ws.schedule()
    frames = [ws.read_frame(f) for f in range(s["first_frame"], s["last_frame"] + 1)]
    return ctx.synthesize(ws.pt_params(ctx.default_params()), frames, ws.read_style(0), scheme=s["scheme"],
                          iterations=s["iterations"], fixup_passes=s["fixup_passes"], want=want)
