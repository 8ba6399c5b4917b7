"""Image files for rendered frames: sRGB 8-bit PNG (after cx_tonemap) and linear PFM.

PNG is written with the standard library only (zlib + the chunk layout of the PNG
specification, RFC 2083): colour type 2 (RGB) or 6 (RGBA), 8 bits, filter type 0 on every row.
read_png decodes any 8-bit non-interlaced PNG (the texture files of the asset loader).  PFM is
the portable float map: "PF", width height, -1.0 (little-endian), rows bottom to top.
"""
from __future__ import annotations

import struct
import zlib

import numpy as np


def _chunk(tag: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def write_png(path: str, img8: np.ndarray):
    """img8: [H, W, 3] (RGB) or [H, W, 4] (RGBA) uint8, row 0 at the top."""
    a = np.ascontiguousarray(img8, np.uint8)
    h, w, c = a.shape
    if c not in (3, 4):
        raise ValueError("expected [H, W, 3] or [H, W, 4] bytes")
    raw = b"".join(b"\x00" + a[y].tobytes() for y in range(h))
    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n")
        f.write(_chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2 if c == 3 else 6, 0, 0, 0)))
        f.write(_chunk(b"IDAT", zlib.compress(raw, 6)))
        f.write(_chunk(b"IEND", b""))


_CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}   # colour type -> samples per pixel


def _paeth_row(line: list, prev: list, bpp: int) -> list:
    cur = list(line)
    for i in range(len(cur)):
        a = cur[i - bpp] if i >= bpp else 0
        b = prev[i]
        c = prev[i - bpp] if i >= bpp else 0
        p = a + b - c
        pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
        pr = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
        cur[i] = (cur[i] + pr) & 255
    return cur


def _average_row(line: list, prev: list, bpp: int) -> list:
    cur = list(line)
    for i in range(len(cur)):
        a = cur[i - bpp] if i >= bpp else 0
        cur[i] = (cur[i] + ((a + prev[i]) >> 1)) & 255
    return cur


def _unfilter(raw: np.ndarray, h: int, stride: int, bpp: int) -> np.ndarray:
    """Reverses the five PNG filter types row by row (PNG specification, section 9)."""
    out = np.zeros((h, stride), np.uint8)
    prev = np.zeros(stride, np.uint8)
    rows = raw.reshape(h, stride + 1)
    for y in range(h):
        ft, line = int(rows[y, 0]), rows[y, 1:]
        if ft == 0:
            cur = line.copy()
        elif ft == 1:      # Sub: running sum of each byte lane, modulo 256
            cur = np.cumsum(line.reshape(-1, bpp), axis=0, dtype=np.uint8).reshape(-1)
        elif ft == 2:      # Up
            cur = line + prev
        elif ft == 3:      # Average
            cur = np.asarray(_average_row(line.tolist(), prev.tolist(), bpp), np.uint8)
        elif ft == 4:      # Paeth
            cur = np.asarray(_paeth_row(line.tolist(), prev.tolist(), bpp), np.uint8)
        else:
            raise ValueError(f"PNG filter type {ft} in row {y}")
        out[y] = cur
        prev = out[y]
    return out


def read_png(path: str, rgba: bool = False) -> np.ndarray:
    """Decodes an 8-bit, non-interlaced PNG of any colour type (grey, grey + alpha, RGB, RGBA,
    palette with optional tRNS) and any filter types.  Returns [H, W, 4] RGBA when `rgba` (grey
    replicated, alpha 255 where the file has none: what stbi_load(..., STBI_rgb_alpha) returns,
    src/Texture.cpp:20), else the [H, W, 3] RGB part."""
    data = open(path, "rb").read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG")
    pos, idat = 8, []
    w = h = ctype = 0
    palette, trns = None, None
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        tag = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        if tag == b"IHDR":
            w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", body[:13])
            if depth != 8 or ctype not in _CHANNELS:
                raise ValueError(f"only 8-bit PNG is read (bit depth {depth}, colour type {ctype})")
            if interlace:
                raise ValueError("interlaced PNG is not read")
        elif tag == b"PLTE":
            palette = np.frombuffer(body, np.uint8).reshape(-1, 3)
        elif tag == b"tRNS":
            trns = np.frombuffer(body, np.uint8)
        elif tag == b"IDAT":
            idat.append(body)
        elif tag == b"IEND":
            break
        pos += 12 + n
    ch = _CHANNELS[ctype]
    raw = np.frombuffer(zlib.decompress(b"".join(idat)), np.uint8)
    if raw.size != h * (1 + w * ch):
        raise ValueError("PNG image data has the wrong size")
    px = _unfilter(raw, h, w * ch, ch).reshape(h, w, ch)
    out = np.full((h, w, 4), 255, np.uint8)
    if ctype == 0:
        out[..., :3] = px
        if trns is not None and trns.size >= 2:
            out[..., 3] = np.where(px[..., 0] == trns[1], 0, 255)
    elif ctype == 4:
        out[..., :3] = px[..., :1]
        out[..., 3] = px[..., 1]
    elif ctype == 2:
        out[..., :3] = px
        if trns is not None and trns.size >= 6:
            key = trns[1::2][:3]
            out[..., 3] = np.where((px == key).all(-1), 0, 255)
    elif ctype == 6:
        out[...] = px
    else:
        if palette is None:
            raise ValueError("palette PNG without PLTE")
        idx = px[..., 0]
        out[..., :3] = palette[idx]
        if trns is not None:
            alpha = np.full(256, 255, np.uint8)
            alpha[:trns.size] = trns
            out[..., 3] = alpha[idx]
    return out if rgba else out[..., :3].copy()


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
