"""Image files for rendered frames: sRGB 8-bit PNG (after cx_tonemap) and linear PFM.

PNG is written with the standard library only (zlib + the chunk layout of the PNG
specification, RFC 2083): colour type 2 (RGB), 8 bits, filter type 0 on every row.  PFM is
the portable float map: "PF", width height, -1.0 (little-endian), rows bottom to top.
"""
from __future__ import annotations

import struct
import zlib

import numpy as np


def write_png(path: str, rgb8: np.ndarray):
    """rgb8: [H, W, 3] uint8, row 0 at the top."""
    a = np.ascontiguousarray(rgb8, np.uint8)
    h, w, c = a.shape
    if c != 3:
        raise ValueError("expected [H, W, 3] bytes")
    raw = b"".join(b"\x00" + a[y].tobytes() for y in range(h))

    def chunk(tag: bytes, data: bytes) -> bytes:
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n")
        f.write(chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)))
        f.write(chunk(b"IDAT", zlib.compress(raw, 6)))
        f.write(chunk(b"IEND", b""))


def read_png(path: str) -> np.ndarray:
    """Reads back what write_png writes (RGB, 8 bits, filter 0)."""
    data = open(path, "rb").read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG")
    pos, idat = 8, b""
    w = h = 0
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        tag = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        if tag == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            if depth != 8 or ctype != 2:
                raise ValueError("only 8-bit RGB")
        elif tag == b"IDAT":
            idat += body
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 3 * w)
    if np.any(raw[:, 0] != 0):
        raise ValueError("only filter type 0")
    return raw[:, 1:].reshape(h, w, 3).copy()


def write_pfm(path: str, rgb: np.ndarray):
    a = np.ascontiguousarray(rgb, np.float32)
    h, w, _ = a.shape
    with open(path, "wb") as f:
        f.write(f"PF\n{w} {h}\n-1.0\n".encode())
        f.write(a[::-1].astype("<f4").tobytes())


def save_frame(ctx, output, path: str):
    """Tonemaps a renderer's device output with cx_tonemap and writes it as PNG."""
    from . import runtime
    write_png(path, runtime.tonemap(ctx, output))
