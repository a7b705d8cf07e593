"""Independent readers/writers for the image files the CLI handles, written
from the OpenEXR file-layout document (magic 0x762f3101, version 2, header
attributes, offset table, chunks holding each line's channels in name order;
scanline chunks of 1 line for NONE/RLE/ZIPS and 16 for ZIP, or tiles with a
`tiles` attribute and the 0x200 version flag; RLE/ZIPS/ZIP data byte-split
into even/odd halves and delta-coded around 128 before the run-length or
zlib stage, and stored raw when that is not smaller) and the PNG
specification (chunks with CRC-32, zlib stream, one filter byte per row).
Used only by the tests to check the C++ writers and to feed the C++
reader."""
import struct
import zlib

import numpy as np

EXR_MAGIC = 20000630


def _attr(name, typ, value):
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(value)) + value


COMPRESSION = {"none": 0, "rle": 1, "zips": 2, "zip": 3}


def _split_and_delta(raw):
    a = np.frombuffer(raw, np.uint8)
    t = np.concatenate([a[0::2], a[1::2]]).astype(np.int32)
    d = t.copy()
    d[1:] = (t[1:] - t[:-1] + 128) & 0xFF
    return d.astype(np.uint8).tobytes()


def _run_length(data):
    """Runs of 3..128 equal bytes as (count - 1, byte); other bytes as
    (-count, count literal bytes), count <= 128."""
    out, i, n = bytearray(), 0, len(data)
    while i < n:
        j = i + 1
        while j < n and j - i < 128 and data[j] == data[i]:
            j += 1
        if j - i >= 3:
            out += bytes([j - i - 1, data[i]])
            i = j
            continue
        j = i
        while j < n and j - i < 128 and not (j + 2 < n and data[j] == data[j + 1] == data[j + 2]):
            j += 1
        out.append((i - j) & 0xFF)
        out += data[i:j]
        i = j
    return bytes(out)


def _pack(raw, compression):
    if compression == "none":
        return raw
    pre = _split_and_delta(raw)
    packed = _run_length(pre) if compression == "rle" else zlib.compress(pre, 6)
    return packed if len(packed) < len(raw) else raw


def _block(channels, names, dt, x0, y0, w, h):
    return b"".join(np.asarray(channels[n][y, x0:x0 + w], dtype=dt).tobytes()
                    for y in range(y0, y0 + h) for n in names)


def write_exr(path, channels, half=False, compression="none", tile=None, mipmap=False, shuffle=None):
    """channels: dict name -> 2-D float array (all the same shape).
    tile = (tw, th) writes a tiled file (one level, or with mipmap=True the
    full chain of rounded-down levels, whose smaller levels readers skip); shuffle = a seed that
    writes the chunks in a random order (line order 'random'), the offset
    table still indexing them by position."""
    names = sorted(channels)
    h, w = channels[names[0]].shape
    ptype = 1 if half else 2
    chlist = b"".join(n.encode() + b"\0" + struct.pack("<iBBBBii", ptype, 0, 0, 0, 0, 1, 1) for n in names) + b"\0"
    box = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    header = struct.pack("<ii", EXR_MAGIC, 2 | (0x200 if tile else 0))
    header += _attr("channels", "chlist", chlist)
    header += _attr("compression", "compression", bytes([COMPRESSION[compression]]))
    header += _attr("dataWindow", "box2i", box)
    header += _attr("displayWindow", "box2i", box)
    header += _attr("lineOrder", "lineOrder", b"\2" if shuffle is not None else b"\0")
    header += _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    header += _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
    header += _attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
    if tile:
        header += _attr("tiles", "tiledesc", struct.pack("<IIB", tile[0], tile[1], 1 if mipmap else 0))
    header += b"\0"
    dt = np.dtype("<f2") if half else np.dtype("<f4")
    chunks = []
    if tile:
        tw, th = tile
        levels = [(0, channels, w, h)]
        # mipmap, rounding down: levels down to 1x1, each tiled like level 0
        for lev in range(1, max(w, h).bit_length() if mipmap else 1):
            lw, lh, k = max(w >> lev, 1), max(h >> lev, 1), 1 << lev
            levels.append((lev, {n: channels[n][:lh * k:k, :lw * k:k] + lev for n in names}, lw, lh))
        for lev, ch, lw, lh in levels:
            for ty in range((lh + th - 1) // th):
                for tx in range((lw + tw - 1) // tw):
                    cw, chh = min(tw, lw - tx * tw), min(th, lh - ty * th)
                    data = _pack(_block(ch, names, dt, tx * tw, ty * th, cw, chh), compression)
                    chunks.append(struct.pack("<iiiii", tx, ty, lev, lev, len(data)) + data)
    else:
        lines = 16 if compression == "zip" else 1
        for y in range(0, h, lines):
            data = _pack(_block(channels, names, dt, 0, y, w, min(lines, h - y)), compression)
            chunks.append(struct.pack("<ii", y, len(data)) + data)
    order = list(range(len(chunks)))
    if shuffle is not None:
        np.random.default_rng(shuffle).shuffle(order)
    off = len(header) + 8 * len(chunks)
    offsets = [0] * len(chunks)
    for k in order:
        offsets[k] = off
        off += len(chunks[k])
    with open(path, "wb") as f:
        f.write(header + struct.pack(f"<{len(chunks)}Q", *offsets) + b"".join(chunks[k] for k in order))


def read_exr(path):
    """Returns dict name -> 2-D float32 array of an uncompressed scanline file."""
    b = open(path, "rb").read()
    magic, version = struct.unpack_from("<ii", b, 0)
    assert magic == EXR_MAGIC and (version & 0xFF) == 2 and not (version & 0x1600), "single-part scanline EXR"
    p = 8
    attrs = {}
    while b[p] != 0:
        e = b.index(b"\0", p)
        name = b[p:e].decode()
        e2 = b.index(b"\0", e + 1)
        size = struct.unpack_from("<i", b, e2 + 1)[0]
        attrs[name] = b[e2 + 5:e2 + 5 + size]
        p = e2 + 5 + size
    p += 1
    assert attrs["compression"] == b"\0", "uncompressed"
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    chans, q, cl = [], 0, attrs["channels"]
    while cl[q] != 0:
        e = cl.index(b"\0", q)
        chans.append((cl[q:e].decode(), struct.unpack_from("<i", cl, e + 1)[0]))
        q = e + 17
    names = [c[0] for c in chans]
    assert names == sorted(names)
    out = {n: np.zeros((h, w), np.float32) for n in names}
    offsets = struct.unpack_from(f"<{h}Q", b, p)
    for y, off in enumerate(offsets):
        ly, size = struct.unpack_from("<ii", b, off)
        assert ly == y0 + y
        pos = off + 8
        for n, t in chans:
            dt = np.dtype("<f2") if t == 1 else np.dtype("<f4")
            out[n][y] = np.frombuffer(b, dtype=dt, count=w, offset=pos).astype(np.float32)
            pos += w * dt.itemsize
        assert pos == off + 8 + size
    return out


def read_png_rgb8(path):
    """Decodes an 8-bit RGB PNG (filter type 0 rows), checking every CRC."""
    b = open(path, "rb").read()
    assert b[:8] == b"\x89PNG\r\n\x1a\n"
    p, idat, ihdr = 8, b"", None
    while p < len(b):
        n, = struct.unpack_from(">I", b, p)
        typ = b[p + 4:p + 8]
        data = b[p + 8:p + 8 + n]
        crc, = struct.unpack_from(">I", b, p + 8 + n)
        assert zlib.crc32(typ + data) & 0xFFFFFFFF == crc, f"CRC of {typ!r}"
        if typ == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", data)
        elif typ == b"IDAT":
            idat += data
        p += 12 + n
        if typ == b"IEND":
            break
    w, h, depth, ctype, _, _, interlace = ihdr
    assert depth == 8 and ctype == 2 and interlace == 0
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 3 * w)
    assert np.all(raw[:, 0] == 0)
    return raw[:, 1:].reshape(h, w, 3)
