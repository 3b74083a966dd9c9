"""Image output: 8-bit RGB PNG files of tone-mapped films (the reference saves its displayed image
through SFML, core/main.cpp:530-540; here the standard library's zlib writes the PNG)."""
import struct
import zlib

import numpy as np


def _chunk(kind, data):
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def write_png(path, rgb8):
    """H x W x 3 uint8 -> an 8-bit RGB PNG (filter 0 on every row)."""
    a = np.ascontiguousarray(rgb8, np.uint8)
    if a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("expected an H x W x 3 uint8 image")
    H, W = a.shape[:2]
    raw = np.concatenate([np.zeros((H, 1), np.uint8), a.reshape(H, W * 3)], axis=1).tobytes()
    png = (b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", W, H, 8, 2, 0, 0, 0))
           + _chunk(b"IDAT", zlib.compress(raw, 6)) + _chunk(b"IEND", b""))
    with open(path, "wb") as f:
        f.write(png)


def read_png(path):
    """Reads back what write_png writes (8-bit RGB, filter 0) and checks every chunk's CRC."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG file")
    pos, idat = 8, b""
    W = H = 0
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        kind, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        if struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0] != zlib.crc32(kind + body) & 0xFFFFFFFF:
            raise ValueError("bad CRC in chunk " + kind.decode())
        if kind == b"IHDR":
            W, H, depth, ctype = struct.unpack(">IIBB", body[:10])
            if depth != 8 or ctype != 2:
                raise ValueError("only 8-bit RGB")
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(H, 1 + W * 3)
    if np.any(raw[:, 0] != 0):
        raise ValueError("only filter 0 rows")
    return raw[:, 1:].reshape(H, W, 3).copy()


# ------------------------------------------------------------------ OpenEXR (core/main.cpp:292-318)
# The reference saves its film as an RGBA half-float EXR (Imf::RgbaOutputFile, alpha 1, the film's
# resolved color in each pixel). Written here as a single-part scanline OpenEXR 2.0 file without
# compression: magic and version, the required header attributes, one offset per scanline, then
# each scanline as (y, byte count, the channels in alphabetical order A, B, G, R, each as
# width little-endian halves). Float-to-half conversion rounds to nearest even, as OpenEXR's half.
_EXR_MAGIC = 20000630
_HALF = 1


def _attr(name, kind, payload):
    return name.encode() + b"\0" + kind.encode() + b"\0" + struct.pack("<i", len(payload)) + payload


def write_exr(path, rgb):
    """H x W x 3 float image -> RGBA half EXR (alpha 1), uncompressed scanlines."""
    a = np.asarray(rgb, np.float32)
    if a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("expected an H x W x 3 float image")
    H, W = a.shape[:2]
    with np.errstate(over="ignore"):  # values beyond the half range become infinities, as in OpenEXR
        chans = {"A": np.ones((H, W), np.float16), "B": a[..., 2].astype(np.float16),
                 "G": a[..., 1].astype(np.float16), "R": a[..., 0].astype(np.float16)}
    chlist = b"".join(n.encode() + b"\0" + struct.pack("<iB3xii", _HALF, 0, 1, 1) for n in sorted(chans)) + b"\0"
    box = struct.pack("<iiii", 0, 0, W - 1, H - 1)
    header = (struct.pack("<ii", _EXR_MAGIC, 2)
              + _attr("channels", "chlist", chlist)
              + _attr("compression", "compression", b"\0")
              + _attr("dataWindow", "box2i", box)
              + _attr("displayWindow", "box2i", box)
              + _attr("lineOrder", "lineOrder", b"\0")
              + _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
              + _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
              + _attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
              + b"\0")
    line_bytes = 4 * W * 2
    first = len(header) + 8 * H
    offsets = struct.pack(f"<{H}Q", *[first + y * (8 + line_bytes) for y in range(H)])
    body = bytearray()
    for y in range(H):
        body += struct.pack("<ii", y, line_bytes)
        for n in sorted(chans):
            body += chans[n][y].astype("<f2").tobytes()
    with open(path, "wb") as f:
        f.write(header + offsets + bytes(body))


def read_exr(path):
    """Reads what write_exr writes (uncompressed scanline half channels) -> dict channel -> H x W
    float32 array, and the header attributes as raw bytes."""
    with open(path, "rb") as f:
        data = f.read()
    magic, version = struct.unpack("<ii", data[:8])
    if magic != _EXR_MAGIC or (version & 0xFF) != 2:
        raise ValueError("not a single-part OpenEXR 2 file")
    pos, attrs = 8, {}
    while data[pos] != 0:
        e = data.index(b"\0", pos); name = data[pos:e].decode(); pos = e + 1
        e = data.index(b"\0", pos); kind = data[pos:e].decode(); pos = e + 1
        n = struct.unpack("<i", data[pos:pos + 4])[0]; pos += 4
        attrs[name] = (kind, data[pos:pos + n]); pos += n
    pos += 1
    if attrs["compression"][1] != b"\0":
        raise ValueError("only uncompressed files")
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    W, H = x1 - x0 + 1, y1 - y0 + 1
    cl, names, q = attrs["channels"][1], [], 0
    while cl[q] != 0:
        e = cl.index(b"\0", q); names.append(cl[q:e].decode()); q = e + 1
        ptype = struct.unpack("<i", cl[q:q + 4])[0]
        if ptype != _HALF:
            raise ValueError("only half channels")
        q += 16
    offsets = struct.unpack(f"<{H}Q", data[pos:pos + 8 * H])
    out = {n: np.zeros((H, W), np.float32) for n in names}
    for off in offsets:
        y, nbytes = struct.unpack("<ii", data[off:off + 8])
        line = np.frombuffer(data[off + 8:off + 8 + nbytes], "<f2").reshape(len(names), W)
        for k, n in enumerate(names):
            out[n][y - y0] = line[k]
    return out, attrs
