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
