"""PIZ compression of OpenEXR scanline chunks (compression 4, 32 lines per
chunk), the codec the reference's ImageIO writes its RGBA images with
(src/PaintTweenCUDA/imageIO.h:131-133, imageIO.cpp:340-371).

Written from the published description of the OpenEXR PIZ codec:

1. The chunk's samples are viewed as 16-bit words (a HALF sample is one
   word, a FLOAT or UINT sample two, low word first) and regrouped
   channel by channel: channel c holds its lines one after another.
2. A 65536-bit map marks the word values that occur (value 0 is never
   marked: decoders always keep it); the values in use are renumbered
   0, 1, 2, ... in ascending order, which bounds the largest value.
3. Each word plane of each channel (a FLOAT channel has two, interleaved)
   goes through a 2D Haar-like wavelet in place, level by level over the
   smaller dimension: the 14-bit form (a sum and a difference in signed
   16-bit arithmetic) when every renumbered value is below 2^14, the
   modulo-2^16 form otherwise.
4. The words are Huffman-coded over 65537 symbols, the last one standing
   for "repeat the previous symbol n more times" (n in 8 bits); codes are
   canonical (per length, numerically increasing with the symbol; longer
   codes take the numerically lower ranges); the table is the code length
   of every symbol between the smallest and largest one used, 6 bits each,
   runs of zero lengths written as one short (59..62: 2..5 zeros) or long
   (63 then 8 bits: 6..261 zeros) code.

Chunk bytes: min and max nonzero bitmap byte (u16 each), those bitmap
bytes, the Huffman stream's byte length (i32), then the stream: symbol
range (u32 min, u32 max, the repeat symbol being max), table length
(u32), bit count (u32), a reserved u32, the table and the bits (MSB first).

Any valid code-length assignment decodes with every PIZ reader; this
writer's Huffman tree breaks frequency ties by symbol order, which may
differ from other writers', so its files need not be byte-identical to
theirs.
"""
import heapq
import struct

import numpy as np

USHORT_RANGE = 1 << 16
BITMAP_SIZE = USHORT_RANGE >> 3
HUF_ENCSIZE = USHORT_RANGE + 1
SHORT_ZEROCODE_RUN = 59
LONG_ZEROCODE_RUN = 63
SHORTEST_LONG_RUN = 2 + LONG_ZEROCODE_RUN - SHORT_ZEROCODE_RUN
LONGEST_LONG_RUN = 255 + SHORTEST_LONG_RUN
MAX_CODE_LENGTH = 58


class PizError(ValueError):
    pass


# ---------------------------------------------------------------- wavelet
def _s16(v):
    """16-bit words read as signed shorts."""
    return ((v + 0x8000) & 0xFFFF) - 0x8000


def _enc(a, b, w14):
    """One forward step of a word pair: (low, high) words."""
    if w14:
        sa, sb = _s16(a), _s16(b)
        return ((sa + sb) >> 1) & 0xFFFF, (sa - sb) & 0xFFFF
    ao = (a + 0x8000) & 0xFFFF
    m = (ao + b) >> 1
    d = ao - b
    m = np.where(d < 0, (m + 0x8000) & 0xFFFF, m)
    return m, d & 0xFFFF


def _dec(lo, hi, w14):
    """The inverse of _enc."""
    if w14:
        sl, sh = _s16(lo), _s16(hi)
        ai = sl + (sh & 1) + (sh >> 1)
        return ai & 0xFFFF, (ai - sh) & 0xFFFF
    b = (lo - (hi >> 1)) & 0xFFFF
    return (hi + b - 0x8000) & 0xFFFF, b


def _levels(nx, ny):
    n = min(nx, ny)
    p = 1
    out = []
    while 2 * p <= n:
        out.append(p)
        p *= 2
    return out


def _level_sets(nx, ny, p):
    p2 = 2 * p
    ys = np.arange(0, ny - p2 + 1, p2)
    xs = np.arange(0, nx - p2 + 1, p2)
    return ys, xs, len(ys) * p2, len(xs) * p2


def wavelet_encode(a, max_value):
    """In-place forward wavelet of a (ny, nx) int64 array of words."""
    w14 = max_value < (1 << 14)
    ny, nx = a.shape
    for p in _levels(nx, ny):
        ys, xs, ye, xe = _level_sets(nx, ny, p)
        Y, X = np.ix_(ys, xs)
        i00, i01 = _enc(a[Y, X], a[Y, X + p], w14)
        i10, i11 = _enc(a[Y + p, X], a[Y + p, X + p], w14)
        a[Y, X], a[Y + p, X] = _enc(i00, i10, w14)
        a[Y, X + p], a[Y + p, X + p] = _enc(i01, i11, w14)
        if nx & p:  # the column left without a partner, row pairs
            lo, hi = _enc(a[ys, xe], a[ys + p, xe], w14)
            a[ys, xe], a[ys + p, xe] = lo, hi
        if ny & p:  # the row left without a partner, column pairs
            lo, hi = _enc(a[ye, xs], a[ye, xs + p], w14)
            a[ye, xs], a[ye, xs + p] = lo, hi


def wavelet_decode(a, max_value):
    """In-place inverse of wavelet_encode."""
    w14 = max_value < (1 << 14)
    ny, nx = a.shape
    for p in reversed(_levels(nx, ny)):
        ys, xs, ye, xe = _level_sets(nx, ny, p)
        Y, X = np.ix_(ys, xs)
        i00, i10 = _dec(a[Y, X], a[Y + p, X], w14)
        i01, i11 = _dec(a[Y, X + p], a[Y + p, X + p], w14)
        a[Y, X], a[Y, X + p] = _dec(i00, i01, w14)
        a[Y + p, X], a[Y + p, X + p] = _dec(i10, i11, w14)
        if nx & p:
            lo, hi = _dec(a[ys, xe], a[ys + p, xe], w14)
            a[ys, xe], a[ys + p, xe] = lo, hi
        if ny & p:
            lo, hi = _dec(a[ye, xs], a[ye, xs + p], w14)
            a[ye, xs], a[ye, xs + p] = lo, hi


# ---------------------------------------------------------------- Huffman
def code_lengths(freq):
    """Huffman code lengths of the symbols with nonzero frequency (a dict
    symbol -> length): the two lightest subtrees are merged until one is
    left (ties go to the subtree holding the smallest symbol), and a
    symbol's length is its leaf's depth."""
    syms = [int(s) for s in np.nonzero(freq)[0]]
    if len(syms) == 1:
        return {syms[0]: 1}
    parent = [-1] * (2 * len(syms) - 1)
    heap = [(int(freq[s]), s, k) for k, s in enumerate(syms)]
    heapq.heapify(heap)
    node = len(syms)
    while len(heap) > 1:
        f1, k1, n1 = heapq.heappop(heap)
        f2, k2, n2 = heapq.heappop(heap)
        parent[n1] = parent[n2] = node
        heapq.heappush(heap, (f1 + f2, min(k1, k2), node))
        node += 1
    depth = [0] * len(parent)
    for n in range(len(parent) - 2, -1, -1):  # parents are created after their children
        depth[n] = depth[parent[n]] + 1
    return {s: depth[k] for k, s in enumerate(syms)}


def canonical_codes(lengths):
    """Canonical codes from an array of code lengths (0 = unused): per
    length, codes increase with the symbol, and the longest codes take the
    lowest values.  Returns an array of codes (unused symbols 0)."""
    lengths = np.asarray(lengths, np.int64)
    count = np.bincount(lengths, minlength=MAX_CODE_LENGTH + 1)
    start = [0] * (MAX_CODE_LENGTH + 1)
    c = 0
    for ln in range(MAX_CODE_LENGTH, 0, -1):
        nxt = (c + int(count[ln])) >> 1
        start[ln] = c
        c = nxt
    codes = np.zeros(len(lengths), np.int64)
    nxt = list(start)
    for s in np.nonzero(lengths)[0]:
        ln = int(lengths[s])
        codes[s] = nxt[ln]
        nxt[ln] += 1
    return codes


def _pack_bits(values, nbits):
    """MSB-first packing of (value, bit count) pairs; returns (bytes, bits)."""
    values = np.asarray(values, np.uint64)
    nbits = np.asarray(nbits, np.int64)
    total = int(nbits.sum())
    if total == 0:
        return b"", 0
    rep_v = np.repeat(values, nbits)
    ends = np.cumsum(nbits)
    pos = np.arange(total) - np.repeat(ends - nbits, nbits)  # bit index inside its field
    shift = (np.repeat(nbits, nbits) - 1 - pos).astype(np.uint64)
    bits = ((rep_v >> shift) & np.uint64(1)).astype(np.uint8)
    return np.packbits(bits).tobytes(), total


def huffman_encode(words):
    """The Huffman stream of an array of 16-bit words (with its 20-byte
    header)."""
    words = np.asarray(words, np.int64)
    freq = np.bincount(words, minlength=HUF_ENCSIZE).astype(np.int64)
    used = np.nonzero(freq)[0]
    im, iM = int(used[0]), int(used[-1]) + 1  # iM: the repeat symbol
    freq[iM] = 1
    lens = np.zeros(HUF_ENCSIZE, np.int64)
    for s, ln in code_lengths(freq).items():
        lens[s] = ln
    if lens.max() > MAX_CODE_LENGTH:
        raise PizError("Huffman code longer than %d bits" % MAX_CODE_LENGTH)
    codes = canonical_codes(lens)
    # table: 6-bit lengths from im to iM, zero runs shortened
    tv, tn = [], []
    s = im
    while s <= iM:
        ln = int(lens[s])
        if ln == 0:
            run = 1
            while s < iM and run < LONGEST_LONG_RUN and lens[s + 1] == 0:
                s += 1
                run += 1
            if run >= SHORTEST_LONG_RUN:
                tv += [LONG_ZEROCODE_RUN, run - SHORTEST_LONG_RUN]
                tn += [6, 8]
                s += 1
                continue
            if run >= 2:
                tv.append(SHORT_ZEROCODE_RUN + run - 2)
                tn.append(6)
                s += 1
                continue
        tv.append(ln)
        tn.append(6)
        s += 1
    table, _ = _pack_bits(tv, tn)
    # data: each run of one symbol in pieces of at most 256, a piece sent as
    # the symbol, the repeat symbol and its extra count when that is shorter
    change = np.nonzero(np.diff(words))[0] + 1
    starts = np.concatenate([[0], change])
    runs = np.diff(np.concatenate([starts, [len(words)]]))
    dv, dn = [], []
    rl, rc = int(lens[iM]), int(codes[iM])
    for st, run in zip(starts, runs):
        sym = int(words[st])
        ln, cd = int(lens[sym]), int(codes[sym])
        while run > 0:
            k = min(run, 256)
            extra = k - 1
            if ln + rl + 8 < ln * extra:
                dv += [cd, rc, extra]
                dn += [ln, rl, 8]
            else:
                dv += [cd] * k
                dn += [ln] * k
            run -= k
    data, nbits = _pack_bits(dv, dn)
    return struct.pack("<IIIII", im, iM, len(table), nbits, 0) + table + data


def huffman_decode(buf, count):
    """`count` 16-bit words from a Huffman stream (huffman_encode's form)."""
    if len(buf) < 20:
        raise PizError("truncated Huffman header")
    im, iM, tlen, nbits, _ = struct.unpack_from("<IIIII", buf, 0)
    if im >= HUF_ENCSIZE or iM >= HUF_ENCSIZE or im > iM:
        raise PizError("bad Huffman symbol range %d..%d" % (im, iM))
    bits = np.unpackbits(np.frombuffer(buf, np.uint8, offset=20))
    pos = 0

    def take(n):
        nonlocal pos
        if pos + n > len(bits):
            raise PizError("Huffman table runs past the stream")
        v = 0
        for b in bits[pos:pos + n]:
            v = (v << 1) | int(b)
        pos += n
        return v

    lens = np.zeros(HUF_ENCSIZE, np.int64)
    s = im
    while s <= iM:
        ln = take(6)
        if ln == LONG_ZEROCODE_RUN:
            run = take(8) + SHORTEST_LONG_RUN
        elif ln >= SHORT_ZEROCODE_RUN:
            run = ln - SHORT_ZEROCODE_RUN + 2
        else:
            lens[s] = ln
            s += 1
            continue
        if s + run > iM + 1:
            raise PizError("Huffman table zero run past the symbol range")
        s += run
    start_bit = 20 * 8 + ((pos + 7) // 8) * 8  # the data starts at a byte after the table
    if tlen and 20 + tlen <= len(buf):
        start_bit = (20 + tlen) * 8
    codes = canonical_codes(lens)
    # a 16-bit look-up for codes of at most 16 bits, a per-length search beyond
    lut_sym = np.full(1 << 16, -1, np.int64)
    lut_len = np.zeros(1 << 16, np.int64)
    longer = {}
    for sym in np.nonzero(lens)[0]:
        ln, cd = int(lens[sym]), int(codes[sym])
        if ln <= 16:
            lo = cd << (16 - ln)
            lut_sym[lo:lo + (1 << (16 - ln))] = sym
            lut_len[lo:lo + (1 << (16 - ln))] = ln
        else:
            longer[(ln, cd)] = int(sym)
    allbits = np.unpackbits(np.frombuffer(buf, np.uint8))
    end = start_bit + nbits
    if end > len(allbits):
        raise PizError("Huffman data shorter than its bit count")
    seg = np.concatenate([allbits[start_bit:end], np.zeros(16, np.uint8)]).astype(np.int64)
    win = np.zeros(nbits + 1, np.int64)
    for k in range(16):
        win |= seg[k:k + nbits + 1] << (15 - k)
    lut_sym_l, lut_len_l, win_l = lut_sym.tolist(), lut_len.tolist(), win.tolist()
    out = []
    p = 0
    prev = None
    while p < nbits:
        w = win_l[p]
        sym = lut_sym_l[w]
        if sym >= 0:
            p += lut_len_l[w]
        else:
            c, ln = 0, 0
            while True:
                if p + ln >= nbits:
                    raise PizError("undecodable Huffman bits")
                c = (c << 1) | int(seg[p + ln])
                ln += 1
                if (ln, c) in longer:
                    sym = longer[(ln, c)]
                    p += ln
                    break
                if ln > MAX_CODE_LENGTH:
                    raise PizError("undecodable Huffman bits")
        if p > nbits:
            raise PizError("Huffman code past the bit count")
        if sym == iM:
            if prev is None or p + 8 > nbits:
                raise PizError("bad repeat in the Huffman data")
            n = win_l[p] >> 8
            p += 8
            out.extend([prev] * n)
        else:
            out.append(sym)
            prev = sym
        if len(out) > count:
            raise PizError("Huffman data decodes to more than %d words" % count)
    if len(out) != count:
        raise PizError("Huffman data decodes to %d words, expected %d" % (len(out), count))
    return np.asarray(out, np.int64)


# ---------------------------------------------------------------- chunks
def _to_planes(raw, sizes, nx, ny):
    """Line-interleaved chunk bytes -> words grouped by channel."""
    w = np.frombuffer(raw, "<u2").astype(np.int64)
    per_line = sum(sizes) * nx
    if len(w) != per_line * ny:
        raise PizError("chunk holds %d words, expected %d" % (len(w), per_line * ny))
    lines = w.reshape(ny, per_line)
    out, off = [], 0
    for s in sizes:
        out.append(lines[:, off:off + s * nx].reshape(-1))
        off += s * nx
    return np.concatenate(out)


def _from_planes(words, sizes, nx, ny):
    parts, off = [], 0
    for s in sizes:
        n = s * nx * ny
        parts.append(words[off:off + n].reshape(ny, s * nx))
        off += n
    return np.concatenate(parts, axis=1).astype("<u2").tobytes()


def _wavelet_all(words, sizes, nx, ny, max_value, fn):
    off = 0
    for s in sizes:
        plane = words[off:off + s * nx * ny].reshape(ny, nx, s)
        for j in range(s):
            a = plane[:, :, j].copy()
            fn(a, max_value)
            plane[:, :, j] = a
        off += s * nx * ny


def compress(raw, sizes, nx, ny):
    """A chunk of `ny` lines of `nx` pixels; sizes: words per sample of each
    channel in file order (HALF 1, FLOAT and UINT 2)."""
    words = _to_planes(raw, sizes, nx, ny)
    present = np.zeros(USHORT_RANGE, bool)
    present[words] = True
    present[0] = False
    bitmap = np.packbits(present, bitorder="little")
    nz = np.nonzero(bitmap)[0]
    lo, hi = (int(nz[0]), int(nz[-1])) if len(nz) else (BITMAP_SIZE - 1, 0)
    keep = present.copy()
    keep[0] = True
    lut = np.zeros(USHORT_RANGE, np.int64)
    lut[keep] = np.arange(int(keep.sum()))
    max_value = int(keep.sum()) - 1
    words = lut[words]
    _wavelet_all(words, sizes, nx, ny, max_value, wavelet_encode)
    huff = huffman_encode(words)
    head = struct.pack("<HH", lo, hi) + (bitmap[lo:hi + 1].tobytes() if lo <= hi else b"")
    return head + struct.pack("<i", len(huff)) + huff


def decompress(data, sizes, nx, ny):
    if len(data) < 4:
        raise PizError("truncated PIZ chunk")
    lo, hi = struct.unpack_from("<HH", data, 0)
    if hi >= BITMAP_SIZE:
        raise PizError("bad PIZ bitmap range")
    pos = 4
    bitmap = np.zeros(BITMAP_SIZE, np.uint8)
    if lo <= hi:
        if pos + hi - lo + 1 > len(data):
            raise PizError("truncated PIZ bitmap")
        bitmap[lo:hi + 1] = np.frombuffer(data, np.uint8, hi - lo + 1, pos)
        pos += hi - lo + 1
    present = np.unpackbits(bitmap, bitorder="little").astype(bool)
    present[0] = True
    lut = np.nonzero(present)[0]
    max_value = len(lut) - 1
    if pos + 4 > len(data):
        raise PizError("truncated PIZ chunk")
    (n,) = struct.unpack_from("<i", data, pos)
    pos += 4
    if n < 0 or pos + n > len(data):
        raise PizError("PIZ Huffman stream runs past the chunk")
    count = sum(sizes) * nx * ny
    words = huffman_decode(data[pos:pos + n], count)
    _wavelet_all(words, sizes, nx, ny, max_value, wavelet_decode)
    if words.max(initial=0) > max_value:
        raise PizError("PIZ wavelet output outside the bitmap's values")
    return _from_planes(lut[words], sizes, nx, ny)
