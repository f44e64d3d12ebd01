"""EXR reader/writer for offset dumps and RGBA images (painttween_amd/exr.py),
the files of ImageIO (src/PaintTweenCUDA/imageIO.cpp:45,340-428).

No OpenEXR library and no EXR file exists in this image or in the reference,
so interoperability with OpenEXR itself is parity unpinned.  The tests pin
the module against byte values worked out by hand from the published file
layout (magic bytes, channel-list records, the ZIP predictor, RLE runs) and
check round trips and the half rounding."""
import struct
import zlib

import numpy as np
import pytest

from painttween_amd import exr, PATCHXF_DTYPE


def test_magic_and_version_bytes(tmp_path):
    p = tmp_path / "a.exr"
    exr.write_exr(p, {"R": (np.zeros((1, 1), np.float32), exr.FLOAT)}, exr.NO_COMPRESSION)
    b = p.read_bytes()
    # 20000630 little-endian, then version 2 with no flags
    assert b[:8] == bytes([0x76, 0x2F, 0x31, 0x01, 2, 0, 0, 0])


def test_uncompressed_file_layout(tmp_path):
    """A 2x1 FLOAT image with channels B and A: the bytes after the header are
    the offset table (one u64 per line) and per line y, size, then each
    channel's samples in name order (A before B)."""
    p = tmp_path / "b.exr"
    a = np.array([[1.0, 2.0]], np.float32)
    bb = np.array([[3.0, 4.0]], np.float32)
    exr.write_exr(p, {"B": (bb, exr.FLOAT), "A": (a, exr.FLOAT)}, exr.NO_COMPRESSION)
    buf = p.read_bytes()
    # channel list: "A\0", type 2, pLinear 0, 3 reserved, sampling 1, 1; then "B"; then "\0"
    chl = b"A\0" + bytes([2, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]) + \
          b"B\0" + bytes([2, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]) + b"\0"
    assert b"channels\0chlist\0" + struct.pack("<i", len(chl)) + chl in buf
    assert b"compression\0compression\0" + bytes([1, 0, 0, 0, 0]) in buf
    assert b"dataWindow\0box2i\0" + struct.pack("<i4i", 16, 0, 0, 1, 0) in buf
    chunk = struct.pack("<ii", 0, 16) + struct.pack("<4f", 1.0, 2.0, 3.0, 4.0)
    assert buf.endswith(chunk)
    table_at = len(buf) - len(chunk) - 8
    assert struct.unpack_from("<Q", buf, table_at)[0] == table_at + 8
    assert buf[table_at - 1] == 0  # end of header


def test_zip_predictor_known_bytes():
    # bytes 1,2,3,4 -> interleaved 1,3,2,4 -> deltas (+128 mod 256) 1,130,127,130
    assert exr._interleave_predict(bytes([1, 2, 3, 4])) == bytes([1, 130, 127, 130])
    assert exr._unpredict_deinterleave(bytes([1, 130, 127, 130])) == bytes([1, 2, 3, 4])
    # odd length: the first half holds the extra even-indexed byte; wrap-around deltas
    # (255, 0, 1 -> 255, 1, 0 -> 255, 1-255+128 = 130 mod 256, 0-1+128 = 127)
    assert exr._interleave_predict(bytes([255, 0, 1])) == bytes([255, 130, 127])
    assert exr._unpredict_deinterleave(bytes([255, 130, 127])) == bytes([255, 0, 1])


def test_zip_chunk_is_zlib_of_predicted_block(tmp_path):
    p = tmp_path / "c.exr"
    v = np.arange(64, dtype=np.float32).reshape(2, 32)
    exr.write_exr(p, {"Y": (v, exr.FLOAT)}, exr.ZIP_COMPRESSION)
    buf = p.read_bytes()
    planes, hdr = exr.read_exr(p)
    assert hdr["compression"] == exr.ZIP_COMPRESSION
    # one 16-line block holds both lines: find it through the offset table
    head_end = buf.index(b"screenWindowWidth\0float\0") + len(b"screenWindowWidth\0float\0") + 8 + 1
    off = struct.unpack_from("<Q", buf, head_end)[0]
    y, size = struct.unpack_from("<ii", buf, off)
    assert y == 0
    raw = zlib.decompress(buf[off + 8:off + 8 + size])
    assert raw == exr._interleave_predict(v.tobytes())
    np.testing.assert_array_equal(planes["Y"], v)


def test_rle_runs():
    # +2 -> next byte 3 times; -3 -> 3 literal bytes
    assert exr._rle_decode(bytes([2, 7, 0xFD, 1, 2, 3]), 6) == bytes([7, 7, 7, 1, 2, 3])
    with pytest.raises(exr.ExrError):
        exr._rle_decode(bytes([2, 7]), 4)


@pytest.mark.parametrize("comp", [exr.NO_COMPRESSION, exr.ZIPS_COMPRESSION, exr.ZIP_COMPRESSION])
def test_offsets_round_trip(tmp_path, comp):
    rng = np.random.default_rng(3)
    h, w = 37, 21  # not a multiple of the 16-line ZIP block
    o = np.zeros((h, w), PATCHXF_DTYPE)
    for f in ("x", "y", "theta", "hysteresis", "scale_u", "scale_v", "lum_y", "lum_a"):
        o[f] = rng.uniform(-50, 300, (h, w)).astype(np.float32)
    o["layer"] = rng.integers(0, 4, (h, w))
    p = tmp_path / "off.exr"
    exr.write_offsets(p, o, comp)
    planes, hdr = exr.read_exr(p)
    assert sorted(planes) == sorted("PatchXF." + c for c in exr.PATCHXF_CHANNELS)
    assert all(t == exr.FLOAT for _, t in hdr["channels"])
    back = exr.read_offsets(p)
    assert back.tobytes() == o.tobytes()


def test_offsets_layer_truncates_like_float_to_int(tmp_path):
    p = tmp_path / "l.exr"
    o = np.zeros((1, 3), PATCHXF_DTYPE)
    exr.write_offsets(p, o)
    planes, _ = exr.read_exr(p)
    planes["PatchXF.layer"][:] = [2.9, -0.5, 1.0]
    exr.write_exr(p, {k: (v, exr.FLOAT) for k, v in planes.items()})
    assert list(exr.read_offsets(p)["layer"][0]) == [2, 0, 1]


def test_rgba_rounds_to_half(tmp_path):
    rng = np.random.default_rng(5)
    img = rng.uniform(0, 1, (9, 13, 4)).astype(np.float32)
    img[0, 0] = [1 + 2 ** -11, 1 + 3 * 2 ** -11, 65520.0, 1e-8]  # ties to even, overflow, subnormal
    p = tmp_path / "rgba.exr"
    with np.errstate(over="ignore"):
        exr.write_rgba(p, img)
        want = img.astype(np.float16).astype(np.float32)
    back = exr.read_rgba(p)
    np.testing.assert_array_equal(back, want)
    assert back[0, 0, 0] == 1.0 and back[0, 0, 1] == 1 + 2 ** -9 and np.isinf(back[0, 0, 2])
    _, hdr = exr.read_exr(p)
    assert [n for n, _ in hdr["channels"]] == ["A", "B", "G", "R"]
    assert all(t == exr.HALF for _, t in hdr["channels"])


def test_rejects_other_files(tmp_path):
    p = tmp_path / "x.exr"
    p.write_bytes(b"\0" * 16)
    with pytest.raises(exr.ExrError):
        exr.read_exr(p)
    exr.write_exr(p, {"R": (np.zeros((2, 2), np.float32), exr.FLOAT)}, exr.NO_COMPRESSION)
    b = bytearray(p.read_bytes())
    b[5] |= 0x02  # tiled flag (bit 9)
    p.write_bytes(bytes(b))
    with pytest.raises(exr.ExrError):
        exr.read_exr(p)
    with pytest.raises(exr.ExrError):
        exr.write_exr(p, {"R": (np.zeros((2, 2), np.float32), exr.FLOAT)}, 6)  # B44: lossy, not written


def _patch_data_window(buf, box):
    key = b"dataWindow\0box2i\0" + struct.pack("<i", 16)
    at = buf.index(key) + len(key)
    return buf[:at] + struct.pack("<4i", *box) + buf[at + 16:]


def test_data_window_away_from_the_origin_reads_unshifted(tmp_path):
    """ImageIO::readImage points its frame buffer at pixels - dx - dy * dw
    (imageIO.cpp:106-117): the array is the data window itself.  A file whose
    data window starts at x = 3 (chunk lines are data-window lines, so only
    the box changes) reads back the same array."""
    p = tmp_path / "w.exr"
    a = np.arange(12, dtype=np.float32).reshape(3, 4)
    exr.write_exr(p, {"R": (a, exr.FLOAT)}, exr.NO_COMPRESSION)
    p.write_bytes(_patch_data_window(p.read_bytes(), (3, 0, 6, 2)))
    planes, hdr = exr.read_exr(p)
    assert (hdr["width"], hdr["height"]) == (4, 3)
    assert planes["R"].tobytes() == a.tobytes()


def test_malformed_chunks_raise_exr_error(tmp_path):
    p = tmp_path / "m.exr"
    a = np.arange(12, dtype=np.float32).reshape(3, 4)
    exr.write_exr(p, {"R": (a, exr.FLOAT)}, exr.NO_COMPRESSION)
    good = p.read_bytes()
    # the first chunk (line 0) sits right after the three u64 offsets
    first = struct.unpack_from("<Q", good, len(good) - 3 * (8 + 16) - 3 * 8)[0]
    for patch in (struct.pack("<ii", 7, 16),      # line outside the data window
                  struct.pack("<ii", 1, 16),      # line 1 twice
                  struct.pack("<ii", 0, 10 ** 6),  # size past the end of the file
                  struct.pack("<ii", 0, 12)):     # fewer bytes than a line holds
        b = bytearray(good)
        b[first:first + 8] = patch
        p.write_bytes(bytes(b))
        with pytest.raises(exr.ExrError):
            exr.read_exr(p)
    b = bytearray(good)
    b[len(good) - 3 * (8 + 16) - 24:len(good) - 3 * (8 + 16) - 16] = struct.pack("<Q", 10 ** 9)
    p.write_bytes(bytes(b))
    with pytest.raises(exr.ExrError):
        exr.read_exr(p)


# ---------------------------------------------------------------- PIZ
from painttween_amd import exr_piz  # noqa: E402


def test_piz_wavelet_steps_worked_by_hand():
    """14-bit step: (a, b) -> ((a + b) >> 1, a - b) in signed 16-bit; a 2x2
    block runs it along rows, then along columns of the row results.
    [[10, 4], [7, 1]]: rows (7, 6), (4, 6); columns (5, 3), (6, 0).  16-bit
    step: a' = (a + 2^15) mod 2^16, m = (a' + b) >> 1, d = a' - b (+2^15 on m
    when d < 0), both mod 2^16: (40000, 1000) -> (4116, 6232)."""
    a = np.array([[10, 4], [7, 1]], np.int64)
    exr_piz.wavelet_encode(a, 100)
    assert a.tolist() == [[5, 6], [3, 0]]
    exr_piz.wavelet_decode(a, 100)
    assert a.tolist() == [[10, 4], [7, 1]]
    lo, hi = exr_piz._enc(np.array([40000]), np.array([1000]), False)
    assert (int(lo[0]), int(hi[0])) == (4116, 6232)
    lo, hi = exr_piz._enc(np.array([1000]), np.array([40000]), False)  # d < 0: m shifted by 2^15
    a2 = (1000 + 32768) & 0xFFFF
    assert (int(lo[0]), int(hi[0])) == ((((a2 + 40000) >> 1) + 32768) & 0xFFFF, (a2 - 40000) & 0xFFFF)
    lo, hi = exr_piz._enc(np.array([3]), np.array([65535]), True)  # -1 as a signed short
    assert (int(lo[0]), int(hi[0])) == ((3 - 1) >> 1, 4)


def test_piz_huffman_streams_worked_by_hand():
    """Header u32 min symbol, max symbol (the repeat symbol, one past the
    largest word), table bytes, bit count, 0; 6-bit code lengths; canonical
    codes with the longer codes numerically lower: lengths 1, 2, 2 give
    codes 1, 00, 01."""
    # [0, 0, 0, 1]: lengths 0:1 1:2 rep:2; table 000001 000010 000010; data 1 1 1 00
    assert exr_piz.huffman_encode(np.array([0, 0, 0, 1])).hex() == \
        "00000000" "02000000" "03000000" "05000000" "00000000" "042080" "e0"
    # [0]*5 + [10]: zeros 1..9 as a long run (63, then 9 - 6 in 8 bits)
    h = exr_piz.huffman_encode(np.array([0] * 5 + [10]))
    assert h[:20] == bytes.fromhex("00000000" "0b000000" "04000000" "07000000" "00000000")
    assert h[20:].hex() == "07f03082" "f8"
    # [0]*5 + [3]: zeros 1..2 as a short run (59 + 2 - 2)
    assert exr_piz.huffman_encode(np.array([0] * 5 + [3]))[20:].hex() == "07b082" "f8"
    # [5]*300 + [7]: the run goes out as 256 (symbol, repeat, 255) and 44 (symbol, repeat, 43)
    h = exr_piz.huffman_encode(np.array([5] * 300 + [7]))
    assert h.hex() == "05000000" "08000000" "03000000" "18000000" "00000000" "040082" "bff4ac"
    for w in ([0, 0, 0, 1], [0] * 5 + [10], [5] * 300 + [7]):
        assert exr_piz.huffman_decode(exr_piz.huffman_encode(np.array(w)), len(w)).tolist() == w
    assert exr_piz.canonical_codes([1, 2, 2]).tolist() == [1, 0, 1]
    assert exr_piz.canonical_codes([2, 0, 3, 3, 1]).tolist() == [1, 0, 0, 1, 1]


def test_piz_round_trips_random_words_and_planes():
    rng = np.random.default_rng(8)
    for trial in range(20):
        n = int(rng.integers(1, 4000))
        w = rng.integers(0, int(rng.integers(2, 70000)), n)
        w[rng.random(n) < 0.4] = w[0]
        w = np.minimum(w, 65535)
        assert (exr_piz.huffman_decode(exr_piz.huffman_encode(w), n) == w).all()
    for mx in (9, 16383, 16384, 65535):
        for ny, nx in ((1, 1), (1, 9), (6, 1), (3, 5), (32, 17), (7, 64), (32, 33), (31, 1000)):
            a = rng.integers(0, mx + 1, (ny, nx))
            b = a.copy()
            exr_piz.wavelet_encode(b, mx)
            exr_piz.wavelet_decode(b, mx)
            assert (a == b).all(), (mx, ny, nx)


def test_piz_files_round_trip(tmp_path):
    """Half RGBA (three 32-line chunks, odd width), FLOAT and UINT channels
    (two words per sample), and constant images; write_rgba writes PIZ."""
    rng = np.random.default_rng(9)
    rgba = rng.uniform(-2, 2, (70, 45, 4)).astype(np.float32)
    rgba[10:40, 5:30] = 0.25  # runs for the repeat symbol
    p = tmp_path / "rgba.exr"
    exr.write_rgba(p, rgba)
    assert b"compression\0compression\0" + struct.pack("<i", 1) + bytes([exr.PIZ_COMPRESSION]) in p.read_bytes()
    assert exr.read_rgba(p).tobytes() == rgba.astype(np.float16).astype(np.float32).tobytes()
    f = rng.uniform(0, 1000, (37, 23)).astype(np.float32)
    u = rng.integers(0, 2 ** 32, (37, 23), dtype=np.uint64).astype(np.uint32)
    exr.write_exr(p, {"a": (f, exr.FLOAT), "b": (u, exr.UINT), "c": (f.astype(np.float16), exr.HALF)},
                  exr.PIZ_COMPRESSION)
    planes, hdr = exr.read_exr(p)
    assert hdr["compression"] == exr.PIZ_COMPRESSION
    assert planes["a"].tobytes() == f.tobytes() and planes["b"].tobytes() == u.tobytes()
    assert planes["c"].tobytes() == f.astype(np.float16).tobytes()
    for v in (0.0, 1.0):
        exr.write_rgba(p, np.full((33, 3, 4), v, np.float32))
        assert np.all(exr.read_rgba(p) == v)


def test_piz_chunk_layout(tmp_path):
    """A chunk: u16 first and last nonzero bitmap byte, those bytes (bit v & 7
    of byte v >> 3 marks word v; word 0 is never marked), i32 Huffman length,
    the Huffman stream.  One 2x2 HALF channel of 1.0 (word 0x3C00) and 0:
    bitmap byte 0x780 holds bit 0."""
    p = tmp_path / "c.exr"
    a = np.array([[1.0, 0.0], [0.0, 1.0]], np.float16)
    exr.write_exr(p, {"R": (a, exr.HALF)}, exr.PIZ_COMPRESSION)
    planes, _ = exr.read_exr(p)
    assert planes["R"].tobytes() == a.tobytes()
    chunk = exr_piz.compress(a.tobytes(), [1], 2, 2)
    lo, hi = struct.unpack_from("<HH", chunk, 0)
    assert (lo, hi) == (0x3C00 >> 3, 0x3C00 >> 3) and chunk[4] == 1
    (n,) = struct.unpack_from("<i", chunk, 5)
    assert n == len(chunk) - 9
    assert exr_piz.decompress(chunk, [1], 2, 2) == a.tobytes()
    z = exr_piz.compress(np.zeros((2, 2), np.float16).tobytes(), [1], 2, 2)
    assert struct.unpack_from("<HH", z, 0) == (exr_piz.BITMAP_SIZE - 1, 0)
    with pytest.raises(exr_piz.PizError):
        exr_piz.decompress(chunk[:-1], [1], 2, 2)
