"""Independent readers/writers for the image files the CLI handles, written
from the OpenEXR file-layout document (magic 0x762f3101, version 2, header
attributes, scanline offset table, uncompressed chunks of one line with the
channels in name order) and the PNG specification (chunks with CRC-32, zlib
stream, one filter byte per row).  Used only by the tests to check the C++
writers and to feed the C++ reader."""
import struct
import zlib

import numpy as np

EXR_MAGIC = 20000630


def _attr(name, typ, value):
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(value)) + value


def write_exr(path, channels, half=False):
    """channels: dict name -> 2-D float array (all the same shape)."""
    names = sorted(channels)
    h, w = channels[names[0]].shape
    ptype = 1 if half else 2
    chlist = b"".join(n.encode() + b"\0" + struct.pack("<iBBBBii", ptype, 0, 0, 0, 0, 1, 1) for n in names) + b"\0"
    box = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    header = struct.pack("<ii", EXR_MAGIC, 2)
    header += _attr("channels", "chlist", chlist)
    header += _attr("compression", "compression", b"\0")
    header += _attr("dataWindow", "box2i", box)
    header += _attr("displayWindow", "box2i", box)
    header += _attr("lineOrder", "lineOrder", b"\0")
    header += _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    header += _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
    header += _attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\0"
    dt = np.dtype("<f2") if half else np.dtype("<f4")
    lines = []
    for y in range(h):
        data = b"".join(np.asarray(channels[n][y], dtype=dt).tobytes() for n in names)
        lines.append(struct.pack("<ii", y, len(data)) + data)
    off = len(header) + 8 * h
    table = b""
    for ln in lines:
        table += struct.pack("<Q", off)
        off += len(ln)
    with open(path, "wb") as f:
        f.write(header + table + b"".join(lines))


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
