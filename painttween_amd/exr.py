"""OpenEXR files of the synthesis path: the offset field as the reference's
9-channel FP32 `PatchXF` layer and RGBA images as half floats
(ImageIO::writeImage / readImage, src/PaintTweenCUDA/imageIO.cpp:45,356-428).

A self-contained reader and writer for single-part scanline files, written
from the published OpenEXR file layout: magic 20000630, version 2, the
attribute header (channels, compression, dataWindow, displayWindow,
lineOrder, pixelAspectRatio, screenWindowCenter, screenWindowWidth), a
line-offset table and one chunk per scanline block.  Channels are stored in
name order, each line holding every channel's samples in turn.
Compressions: NONE, RLE, ZIPS (one line per block) and ZIP (16 lines per
block), the latter two as zlib streams over the byte-interleaved,
delta-predicted block, and PIZ (32 lines per block, wavelet + Huffman,
painttween_amd/exr_piz.py), the codec of the reference's half RGBA images
(ImageIO::writeImage's default, imageIO.h:131-133); its float layers use
the OpenEXR header's default, ZIP.  The lossy codecs (PXR24, B44, DWA) are
neither read nor written.
"""
import struct
import zlib

import numpy as np

from . import _ffi
from . import exr_piz

MAGIC = 20000630
NO_COMPRESSION, RLE_COMPRESSION, ZIPS_COMPRESSION, ZIP_COMPRESSION, PIZ_COMPRESSION = 0, 1, 2, 3, 4
UINT, HALF, FLOAT = 0, 1, 2
_LINES = {NO_COMPRESSION: 1, RLE_COMPRESSION: 1, ZIPS_COMPRESSION: 1, ZIP_COMPRESSION: 16, PIZ_COMPRESSION: 32}
_WORDS = {UINT: 2, HALF: 1, FLOAT: 2}  # 16-bit words per sample (PIZ)
_NP = {UINT: np.dtype("<u4"), HALF: np.dtype("<f2"), FLOAT: np.dtype("<f4")}

#: channel names of the reference's PatchXF layer (imageIO.cpp:45), in the
#: order of the pt_patchxf fields
PATCHXF_CHANNELS = ("x", "y", "theta", "hyst", "layer", "lumShift_y", "lumShift_a", "scale_u", "scale_v")
_PATCHXF_FIELDS = ("x", "y", "theta", "hysteresis", "layer", "lum_y", "lum_a", "scale_u", "scale_v")


class ExrError(ValueError):
    pass


# ---------------------------------------------------------------- codecs
def _interleave_predict(raw):
    """ZIP preprocessing: even bytes then odd bytes, then byte deltas + 128."""
    b = np.frombuffer(raw, np.uint8)
    t = np.concatenate([b[0::2], b[1::2]]).astype(np.int16)
    d = t.copy()
    d[1:] = (t[1:] - t[:-1] + 128) & 0xFF
    return d.astype(np.uint8).tobytes()


def _unpredict_deinterleave(data):
    t = np.frombuffer(data, np.uint8).astype(np.int64)
    t[1:] -= 128
    t = (np.cumsum(t) & 0xFF).astype(np.uint8)
    n = len(t)
    half = (n + 1) // 2
    out = np.empty(n, np.uint8)
    out[0::2] = t[:half]
    out[1::2] = t[half:]
    return out.tobytes()


def _rle_decode(data, size):
    out = bytearray()
    i = 0
    while i < len(data) and len(out) < size:
        count = struct.unpack_from("b", data, i)[0]
        i += 1
        if count < 0:
            out += data[i:i - count]
            i -= count
        else:
            out += bytes([data[i]]) * (count + 1)
            i += 1
    if len(out) != size:
        raise ExrError("RLE block decodes to %d bytes, expected %d" % (len(out), size))
    return bytes(out)


def _compress(raw, compression):
    if compression == NO_COMPRESSION:
        return raw
    if compression in (ZIP_COMPRESSION, ZIPS_COMPRESSION):
        packed = zlib.compress(_interleave_predict(raw))
        return packed if len(packed) < len(raw) else raw
    raise ExrError("writing compression %d is not supported" % compression)


def _decompress(data, size, compression):
    if len(data) == size:  # stored uncompressed (any codec may do so)
        return data
    if compression in (ZIP_COMPRESSION, ZIPS_COMPRESSION):
        try:
            return _unpredict_deinterleave(zlib.decompress(data))
        except zlib.error as e:
            raise ExrError("corrupt ZIP chunk: %s" % e)
    if compression == RLE_COMPRESSION:
        return _unpredict_deinterleave(_rle_decode(data, size))
    raise ExrError("compression %d is not supported for reading" % compression)


# ---------------------------------------------------------------- header
def _attr(name, typ, payload):
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(payload)) + payload


def _header(width, height, chans, compression):
    cl = b""
    for name, ptype in chans:
        cl += name.encode() + b"\0" + struct.pack("<iB3xii", ptype, 0, 1, 1)
    cl += b"\0"
    box = struct.pack("<iiii", 0, 0, width - 1, height - 1)
    return (struct.pack("<ii", MAGIC, 2) +
            _attr("channels", "chlist", cl) +
            _attr("compression", "compression", struct.pack("<B", compression)) +
            _attr("dataWindow", "box2i", box) +
            _attr("displayWindow", "box2i", box) +
            _attr("lineOrder", "lineOrder", struct.pack("<B", 0)) +
            _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0)) +
            _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0)) +
            _attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) +
            b"\0")


def _cstr(buf, pos):
    end = buf.index(b"\0", pos)
    return buf[pos:end].decode(), end + 1


def _parse_header(buf):
    magic, version = struct.unpack_from("<ii", buf, 0)
    if magic != MAGIC:
        raise ExrError("not an OpenEXR file")
    if version & 0xFF != 2 or version & 0x1A00:  # tiled, deep or multi-part
        raise ExrError("only single-part scanline files are supported (version field %#x)" % version)
    pos = 8
    attrs = {}
    while buf[pos] != 0:
        name, pos = _cstr(buf, pos)
        typ, pos = _cstr(buf, pos)
        size = struct.unpack_from("<i", buf, pos)[0]
        pos += 4
        attrs[name] = (typ, buf[pos:pos + size])
        pos += size
    pos += 1
    chans = []
    cl = attrs["channels"][1]
    p = 0
    while cl[p] != 0:
        name, p = _cstr(cl, p)
        ptype, _, xs, ys = struct.unpack_from("<iB3xii", cl, p)
        p += 16
        if xs != 1 or ys != 1:
            raise ExrError("subsampled channel %s is not supported" % name)
        chans.append((name, ptype))
    xmin, ymin, xmax, ymax = struct.unpack("<iiii", attrs["dataWindow"][1])
    compression = attrs["compression"][1][0]
    return dict(channels=chans, compression=compression, width=xmax - xmin + 1, height=ymax - ymin + 1,
                ymin=ymin, attributes=attrs), pos


# ---------------------------------------------------------------- files
def write_exr(path, channels, compression=ZIP_COMPRESSION):
    """channels: {name: (array (H, W), HALF or FLOAT or UINT)}; all the same size."""
    if not channels:
        raise ExrError("no channels")
    if compression not in (NO_COMPRESSION, ZIPS_COMPRESSION, ZIP_COMPRESSION, PIZ_COMPRESSION):
        raise ExrError("writing compression %d is not supported" % compression)
    names = sorted(channels)
    h, w = np.asarray(channels[names[0]][0]).shape
    chans = []
    planes = []
    for n in names:
        a, ptype = channels[n]
        a = np.ascontiguousarray(a, _NP[ptype])
        if a.shape != (h, w):
            raise ExrError("channel %s is %s, expected %s" % (n, a.shape, (h, w)))
        chans.append((n, ptype))
        planes.append(a)
    lines = _LINES[compression]
    head = _header(w, h, chans, compression)
    nblocks = (h + lines - 1) // lines
    chunks = []
    for bi in range(nblocks):
        y0, y1 = bi * lines, min(h, bi * lines + lines)
        raw = b"".join(p[y].tobytes() for y in range(y0, y1) for p in planes)
        if compression == PIZ_COMPRESSION:
            data = exr_piz.compress(raw, [_WORDS[t] for _, t in chans], w, y1 - y0)
            if len(data) >= len(raw):
                data = raw  # stored as is when the codec does not shrink it
        else:
            data = _compress(raw, compression)
        chunks.append(struct.pack("<ii", y0, len(data)) + data)
    off = len(head) + 8 * nblocks
    table = b""
    for c in chunks:
        table += struct.pack("<Q", off)
        off += len(c)
    with open(path, "wb") as f:
        f.write(head + table + b"".join(chunks))


def read_exr(path):
    """Returns ({name: array (H, W)}, header dict).  The arrays hold the data
    window, row 0 / column 0 at its minimum corner, as ImageIO::readImage
    does with its frame buffer at pixels - dx - dy * dw (imageIO.cpp:106-117),
    so a data window away from the origin is read without a shift.  Malformed
    chunk tables or chunks raise ExrError."""
    with open(path, "rb") as f:
        buf = f.read()
    hdr, pos = _parse_header(buf)
    w, h, comp = hdr["width"], hdr["height"], hdr["compression"]
    if w < 1 or h < 1:
        raise ExrError("empty data window")
    if comp not in _LINES:
        raise ExrError("compression %d is not supported for reading" % comp)
    lines = _LINES[comp]
    nblocks = (h + lines - 1) // lines
    if pos + 8 * nblocks > len(buf):
        raise ExrError("truncated offset table")
    offsets = struct.unpack_from("<%dQ" % nblocks, buf, pos)
    chans = hdr["channels"]
    planes = {n: np.empty((h, w), _NP[t]) for n, t in chans}
    seen = set()
    for off in offsets:
        if off + 8 > len(buf):
            raise ExrError("chunk offset %d beyond the file" % off)
        y, size = struct.unpack_from("<ii", buf, off)
        y -= hdr["ymin"]
        if y < 0 or y >= h or y % lines or y in seen:
            raise ExrError("chunk at line %d outside the data window or repeated" % (y + hdr["ymin"]))
        if size < 0 or off + 8 + size > len(buf):
            raise ExrError("chunk of %d bytes at %d runs past the file" % (size, off))
        seen.add(y)
        n = min(lines, h - y)
        expect = sum(_NP[t].itemsize for _, t in chans) * w * n
        if comp == PIZ_COMPRESSION and size != expect:
            try:
                raw = exr_piz.decompress(buf[off + 8:off + 8 + size], [_WORDS[t] for _, t in chans], w, n)
            except exr_piz.PizError as e:
                raise ExrError("PIZ chunk at line %d: %s" % (y + hdr["ymin"], e))
        else:
            raw = _decompress(buf[off + 8:off + 8 + size], expect, comp)
        if len(raw) != expect:
            raise ExrError("chunk at line %d holds %d bytes, expected %d" % (y + hdr["ymin"], len(raw), expect))
        p = 0
        for yy in range(y, y + n):
            for name, t in chans:
                k = _NP[t].itemsize * w
                planes[name][yy] = np.frombuffer(raw, _NP[t], w, p)
                p += k
    return planes, hdr


def write_offsets(path, offsets, compression=ZIP_COMPRESSION):
    """The offset field (pt_patchxf array, (H, W)) as the reference's FP32
    PatchXF layer (ImageIO::writeImage(QVector<PatchXF>), imageIO.cpp:373-390)."""
    off = np.asarray(offsets)
    ch = {"PatchXF." + c: (off[f].astype(np.float32), FLOAT) for c, f in zip(PATCHXF_CHANNELS, _PATCHXF_FIELDS)}
    write_exr(path, ch, compression)


def read_offsets(path):
    """A PatchXF layer file back into a pt_patchxf array (imageIO.cpp:160-190)."""
    planes, hdr = read_exr(path)
    out = np.zeros((hdr["height"], hdr["width"]), _ffi.PATCHXF_DTYPE)
    for c, f in zip(PATCHXF_CHANNELS, _PATCHXF_FIELDS):
        key = "PatchXF." + c
        if key not in planes:
            raise ExrError("%s has no channel %s" % (path, key))
        v = planes[key].astype(np.float32)
        out[f] = v.astype(np.int32) if f == "layer" else v
    return out


def write_rgba(path, rgba, compression=PIZ_COMPRESSION):
    """An RGBA float image as half-float R, G, B, A channels with the
    reference's PIZ compression (it rounds canvas and residual images to
    half, imageIO.cpp:356-371)."""
    a = np.asarray(rgba, np.float32)
    with np.errstate(over="ignore"):  # beyond 65504 rounds to infinity, as OpenEXR's half does
        write_exr(path, {c: (a[..., k].astype(np.float16), HALF) for k, c in enumerate("RGBA")},
                  compression)


def read_rgba(path):
    planes, _ = read_exr(path)
    return np.stack([planes[c].astype(np.float32) for c in "RGBA"], axis=-1)
