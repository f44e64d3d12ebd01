"""Working-set XML (painttween_amd/working_set.py) against the reference's
rules: NLSynthesizer::readWorkingSet (G/nlSynthesizer.cpp:101-161), NLShot::setup
(G/nlShot.cpp:55-194), NLStyles::load (G/nlStyles.cpp:47-82), NLParameters::load
(G/nlParameters.cpp:146-191) and SynthesisProcessor::updateParameters
(C/synthesisProcessor.cpp:167-262).  The reference ships no working set, so
the files here are synthetic ones laid out as those functions read them."""
import os

import numpy as np
import pytest

from painttween_amd import _ffi, exr, working_set as wsm
from painttween_amd import workloads as W

XML = """<doc creator="me@host">
  <anim>
    <unit name="u1"/>
    <shot name="shotA"/>
    <shotDir path="shots"/>
    <workingDir path="out/../work"/>
    <elements inputElem=".color." orientationElem="orient" distTransElem="dt" velBElem="vb" velFElem="vf"
              ribbonBElem="rb[xx]" ribbonFElem="rf[xx]"/>
    <fileAccess checkFarm="0" checkFiles="1"/>
  </anim>
  <styles>
    <style>
      <styleInput path="style/in.exr"/>
      <styleOutput path="style/out.exr"/>
      <styleOrientation path="style/or.exr"/>
    </style>
  </styles>
  <params>
    <levels value="2"/>
    <iterations value="1"/>
    <residualWindowSize value="2"/>
    <advectionJitter value="5.3"/>
    <distanceTransformWeight value="0.5"/>
    <orientationEnabled value="1"/>
    <transparencyAllowed value="true"/>
    <offsetsHistogramSlope value0="0.25" value1="9"/>
    <synthesisScheme value="0"/>
    <firstFrame value="1"/>
    <lastFrame value="2"/>
    <maxDistortion value="bad"/>
  </params>
</doc>
"""

H, Wd, SH, SW = 24, 28, 20, 22


def _layer(path, layer, arr, channels):
    exr.write_exr(path, {layer + "." + c: (arr[..., k], exr.FLOAT) for k, c in enumerate(channels)})


def _make(tmp_path, xml=XML, frames=(1, 2), vel_f_frames=None):
    d = tmp_path
    (d / "shots").mkdir()
    (d / "style").mkdir()
    st = W.style_pair(SW, SH, seed=7, guides=True)
    _layer(d / "style/in.exr", "style_input", st["base"], ("red", "green", "blue", "alpha"))
    _layer(d / "style/out.exr", "style_output", st["output"], ("red", "green", "blue", "alpha"))
    _layer(d / "style/or.exr", "orientation", st["orientation"], ("Ix2", "IxIy", "Iy2"))
    seq = W.animation(Wd, H, len(frames), seed=8)
    written = {}
    for i, f in enumerate(frames):
        fr = seq[i]
        c = fr["color"].copy()
        exr.write_exr(d / ("shots/shotA.color.%d.exr" % f),
                      {n: (c[..., k], exr.FLOAT) for k, n in enumerate("RGB")})  # no alpha channel
        _layer(d / ("shots/shotA.orient.%d.exr" % f), "orientation", fr["orientation"][..., :2], ("Ix2", "IxIy"))
        _layer(d / ("shots/shotA.dt.%d.exr" % f), "dist_trans", fr["dist_trans"][..., None], ("length",))
        _layer(d / ("shots/shotA.vb.%d.exr" % f), "vel_b_1", fr["vel_b"], ("x", "y"))
        if vel_f_frames is None or f in vel_f_frames:
            _layer(d / ("shots/shotA.vf.%d.exr" % f), "vel_f_1", fr["vel_f"], ("x", "y"))
        written[f] = fr
    (d / "ws.xml").write_text(xml)
    return d / "ws.xml", st, written


def test_load_paths_elements_and_params(tmp_path):
    path, _, _ = _make(tmp_path)
    ws = wsm.load(path)
    assert ws.creator == "me@host" and ws.unit == "u1" and ws.shot == "shotA"
    assert ws.shot_dir == str(tmp_path / "shots") and ws.working_dir == str(tmp_path / "work")
    assert ws.elements["input"] == "color"  # leading and trailing dots removed
    assert ws.input_paths["input"] == str(tmp_path / "shots/shotA") + ".color.%1.exr"
    assert ws.input_path("vel_b", 2) == str(tmp_path / "shots/shotA.vb.2.exr")
    assert ws.styles == [{"input": str(tmp_path / "style/in.exr"), "output": str(tmp_path / "style/out.exr"),
                          "orientation": str(tmp_path / "style/or.exr")}]
    g = ws.get
    assert (g("levels"), g("iterations"), g("residualWindowSize")) == (2, 1, 2)
    assert g("orientationEnabled") is True
    assert g("transparencyAllowed") is False      # "true".toInt() == 0
    assert g("maxDistortion") == 0.0              # unparsable text reads as 0
    assert g("temporalCoherenceWeight") == 0.2    # default kept
    assert g("edgeBlendingWeight") == 0           # no default: unknown reads as 0
    assert ws.histogram_slopes == [0.25]          # resized to the number of styles
    assert ws.histogram_thresholds == [0.0] * 5   # absent: left at the defaults
    assert ws.ribbons_b == [] and ws.ribbons_f == []  # no ribbon files


def test_pt_params_follow_update_parameters(tmp_path):
    path, _, _ = _make(tmp_path)
    ws = wsm.load(path)
    p = ws.pt_params()
    assert p.levels == 2 and p.residual_window_size == 2 and p.num_styles == 1
    assert p.dist_trans_weight == 0.5 and p.transparency_ok == 0 and p.orientation_enabled == 1
    # float("5.3") to float, times the double 0.01, back to float
    assert p.advection_jitter == np.float32(float(np.float32(5.3)) * 0.01)
    assert p.advection_jitter != np.float32(np.float32(5.3) * np.float32(0.01))
    assert p.offsets_histogram_slope[0] == 0.25 and p.offsets_histogram_slope[1] == 0.0
    assert p.max_distortion == 0.0 and p.dist_trans_cull == 50 and p.num_frames == 1
    d = wsm.WorkingSet(path="x").pt_params()
    assert d.temporal_coherence_weight == np.float32(0.2) and d.levels == 4 and d.scatter_samples_count == 25


def test_read_frames_and_style(tmp_path):
    path, st, fr = _make(tmp_path)
    ws = wsm.load(path)
    f = ws.read_frame(2)
    c = fr[2]["color"]
    # RGBA input: rounded to half, missing alpha reads 1
    np.testing.assert_array_equal(f["color"][..., :3], c[..., :3].astype(np.float16).astype(np.float32))
    assert np.all(f["color"][..., 3] == 1.0)
    # FP32 layers: exact, a missing channel reads 0
    np.testing.assert_array_equal(f["orientation"][..., :2], fr[2]["orientation"][..., :2])
    assert np.all(f["orientation"][..., 2] == 0.0)
    np.testing.assert_array_equal(f["dist_trans"], fr[2]["dist_trans"])
    np.testing.assert_array_equal(f["vel_f"], fr[2]["vel_f"])
    assert f["scale"] is None  # no scale element in the shot
    s = ws.read_style(0)
    np.testing.assert_array_equal(s["base"], st["base"])
    np.testing.assert_array_equal(s["output"], st["output"])
    np.testing.assert_array_equal(s["orientation"], st["orientation"])
    assert s["dist_trans"] is None


def test_missing_frames(tmp_path):
    path, _, _ = _make(tmp_path, vel_f_frames=(1,))
    ws = wsm.load(path)
    assert "vel_f" not in ws.input_paths and "vel_b" in ws.input_paths  # dropped, not fatal
    os.remove(tmp_path / "shots/shotA.color.2.exr")
    with pytest.raises(wsm.WorkingSetError, match="cannot load input"):
        wsm.load(path)
    assert "input" in wsm.load(path, check_files=False).input_paths


def test_required_tags(tmp_path):
    path, _, _ = _make(tmp_path)
    for cut, msg in (("<unit name=\"u1\"/>", "unit"), ("<styleInput path=\"style/in.exr\"/>", "input style"),
                     ("<fileAccess checkFarm=\"0\" checkFiles=\"1\"/>", "fileAccess")):
        path.write_text(XML.replace(cut, ""))
        with pytest.raises(wsm.WorkingSetError, match=msg):
            wsm.load(path)


def test_schedule_rejects_unimplemented_settings(tmp_path):
    path, _, _ = _make(tmp_path)
    ws = wsm.load(path)
    assert ws.schedule() == dict(scheme=_ffi.PT_S_FMBM, iterations=1, fixup_passes=1, first_frame=1, last_frame=2)
    ws1 = wsm.load(path)
    ws1.params["firstPassDirection"] = 1
    assert ws1.pt_params().first_pass_direction == 1
    ws1.params["maxLevel"] = 2
    assert ws1.pt_params().max_level == 2 and ws1.schedule()["scheme"] == _ffi.PT_S_FMBM
    for k, v in (("keyFrameInterpolation", True),):
        ws2 = wsm.load(path)
        ws2.params[k] = v
        with pytest.raises(wsm.WorkingSetError, match=k):
            ws2.schedule()


def test_ribbons_and_farm_layout(tmp_path):
    path, _, _ = _make(tmp_path)
    # ribbon names are relative to the working directory of the process (nlShot.cpp:167-180)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        for step in ("01", "02"):
            for f in (1, 2):
                for side in ("rb", "rf"):
                    (tmp_path / ("%s%s.%d.exr" % (side, step, f))).write_bytes(b"")
        ws = wsm.load(path)
        assert ws.ribbons_b == ["rb01.%1.exr", "rb02.%1.exr"] and ws.ribbons_f == ["rf01.%1.exr", "rf02.%1.exr"]
    finally:
        os.chdir(cwd)
    farm = XML.replace('checkFarm="0"', 'checkFarm="1"')
    path.write_text(farm)
    ws = wsm.load(path, check_files=False)
    assert ws.from_farm
    assert ws.input_paths["vel_b"] == str(tmp_path / "shots") + "/vb/shotA.vb.%1.exr"
    assert ws.ribbons_b[0] == str(tmp_path / "shots") + "/rb01/shotA.rb01.%1.exr"


def test_save_load_round_trip(tmp_path):
    path, _, _ = _make(tmp_path)
    ws = wsm.load(path)
    out = tmp_path / "again.xml"
    ws.save(out)
    ws2 = wsm.load(out)
    assert ws2.params == ws.params
    assert ws2.styles == ws.styles and ws2.elements == ws.elements and ws2.input_paths == ws.input_paths
    assert ws2.histogram_slopes == ws.histogram_slopes
    assert (ws2.unit, ws2.shot, ws2.shot_dir, ws2.working_dir) == (ws.unit, ws.shot, ws.shot_dir, ws.working_dir)


def test_synthesize_on_the_oracle(tmp_path):
    """The glue end to end on the CPU restatement: the loaded working set
    gives what Context.synthesize gives with the same arrays by hand."""
    from tests.oracle_ctx import OracleContext
    # the histogram slope 0.25 stays in: with the reference's all-zero histogram
    # and threshold 0 its term is 0, so the by-hand parameters leave it out
    path, st, fr = _make(tmp_path)
    ws = wsm.load(path)
    with OracleContext() as o:
        off, canvas, res = wsm.synthesize(o, ws)
        frames = [ws.read_frame(f) for f in (1, 2)]
        p = o.default_params(levels=2, residual_window_size=2, dist_trans_weight=0.5, max_distortion=0.0,
                             transparency_ok=0, advection_jitter=float(np.float32(5.3)) * 0.01, num_frames=1)
        off2, canvas2, res2 = o.synthesize(p, frames, ws.read_style(0), scheme=_ffi.PT_S_FMBM, iterations=1,
                                           fixup_passes=1)
    assert off.tobytes() == off2.tobytes() and canvas.tobytes() == canvas2.tobytes()
    assert res.tobytes() == res2.tobytes()
    assert off.shape == (2, H, Wd)


@pytest.mark.gpu
def test_synthesize_gpu_matches_oracle(tmp_path):
    """A working set rendered through the HIP library equals the oracle's render."""
    from painttween_amd import Context
    from tests.oracle_ctx import OracleContext
    path, _, _ = _make(tmp_path)
    ws = wsm.load(path)
    print("working set", path, "frames 1-2, 28x24, FMBM", flush=True)
    with Context.gpu(0) as g, OracleContext() as o:
        got, want = wsm.synthesize(g, ws), wsm.synthesize(o, ws)
    for a, b in zip(got, want):
        assert a.tobytes() == b.tobytes()


def test_object_id_elements(tmp_path):
    """The shot's and the style's ID images become r + 255 g + 255^2 b codes
    and reach the object-ID guides."""
    xml = XML.replace('velFElem="vf"', 'velFElem="vf" idMergedElem="ids"').replace(
        '<styleOrientation path="style/or.exr"/>', '<styleOrientation path="style/or.exr"/><styleObjectIDs path="style/id.exr"/>')
    path, _, _ = _make(tmp_path, xml=xml)
    rng = np.random.default_rng(9)
    q = rng.integers(0, 256, (H, Wd, 3))
    for f in (1, 2):
        _layer(tmp_path / ("shots/shotA.ids.%d.exr" % f), "id", (q / 255.0).astype(np.float32), ("red", "green", "blue"))
    qs = rng.integers(0, 256, (SH, SW, 3))
    _layer(tmp_path / "style/id.exr", "id", (qs / 255.0).astype(np.float32), ("red", "green", "blue"))
    ws = wsm.load(path)
    want = (q[..., 0] + 255 * q[..., 1] + 255 * 255 * q[..., 2]).astype(np.int32)
    np.testing.assert_array_equal(ws.read_frame(1)["object_id"], want)
    want_s = (qs[..., 0] + 255 * qs[..., 1] + 255 * 255 * qs[..., 2]).astype(np.int32)
    np.testing.assert_array_equal(ws.read_style(0)["object_ids"], want_s)
    # the two rounding rules differ only below zero: -0.5 rounds to 0 (half up) and -1 (half away)
    half = np.array([[[-0.5 / 255, 0, 0]]], np.float32)
    v = float(half[0, 0, 0] * np.float32(255))
    if v == -0.5:
        assert wsm.ids_from_rgb(half, style=False)[0, 0] == 0 and wsm.ids_from_rgb(half, style=True)[0, 0] == -1


def test_style_id_rounding_is_exact_just_below_a_half():
    """round() of the float product v*255 rounds half away from zero; a float
    just below .5 (0.49999997) must give 0, which a float32 v*255 + 0.5 would
    round up to 1 (G/nlDataAccess.cpp:390).  Shot inputs keep floorf(x + 0.5f)
    in float, where the same value does round to 1 (imageIO.cpp:298-317)."""
    below = np.nextafter(np.float32(0.5), np.float32(0))
    for p, want_style in ((below, 0), (np.float32(0.5), 1), (-below, 0), (np.float32(-0.5), -1),
                          (np.float32(2 ** 23 - 0.5), 2 ** 23)):
        v = np.float32(p) / np.float32(255)
        if v * np.float32(255) != np.float32(p):
            continue  # p is not a product the division reaches exactly
        rgb = np.array([[[v, 0, 0]]], np.float32)
        assert wsm.ids_from_rgb(rgb, style=True)[0, 0] == want_style, p
    rgb = np.array([[[below / np.float32(255), 0, 0]]], np.float32)
    if rgb[0, 0, 0] * np.float32(255) == below:
        assert wsm.ids_from_rgb(rgb, style=False)[0, 0] == 1  # floorf(0.49999997f + 0.5f) = 1 in float
