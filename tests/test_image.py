"""The PNG reader of the asset loader (textures are read as stbi_load(..., STBI_rgb_alpha) reads
them, src/Texture.cpp:20) against Pillow's decoder, on files that use every colour type of
8-bit PNG and every one of the five row filters."""
import struct
import zlib

import numpy as np
import pytest

from crisoptix_amd import image

PIL = pytest.importorskip("PIL.Image")


def _filtered_png(path, img, ctype, filters):
    """A PNG with the given filter type per row (PNG specification, section 9), written here so
    that every filter occurs; Pillow's decoding of it checks the encoding."""
    h, w, ch = img.shape
    bpp = ch
    stride = w * ch
    rows = img.reshape(h, stride).astype(np.int64)
    raw = bytearray()
    prev = np.zeros(stride, np.int64)
    for y in range(h):
        ft = filters[y % len(filters)]
        cur = rows[y]
        left = np.concatenate([np.zeros(bpp, np.int64), cur[:-bpp]])
        upleft = np.concatenate([np.zeros(bpp, np.int64), prev[:-bpp]])
        if ft == 0:
            f = cur
        elif ft == 1:
            f = cur - left
        elif ft == 2:
            f = cur - prev
        elif ft == 3:
            f = cur - (left + prev) // 2
        else:
            p = left + prev - upleft
            pa, pb, pc = np.abs(p - left), np.abs(p - prev), np.abs(p - upleft)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, prev, upleft))
            f = cur - pred
        raw.append(ft)
        raw += (f % 256).astype(np.uint8).tobytes()
        prev = cur

    def chunk(tag, data):
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    with open(path, "wb") as fh:
        fh.write(b"\x89PNG\r\n\x1a\n")
        fh.write(chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0)))
        fh.write(chunk(b"IDAT", zlib.compress(bytes(raw))))
        fh.write(chunk(b"IEND", b""))


@pytest.mark.parametrize("ctype,ch", [(0, 1), (2, 3), (4, 2), (6, 4)])
def test_every_filter_and_colour_type(tmp_path, ctype, ch):
    rng = np.random.default_rng(ctype)
    yy, xx = np.mgrid[0:23, 0:31]
    img = ((rng.integers(0, 256, (23, 31, ch)) // 4) + (7 * xx + 3 * yy)[..., None]).astype(np.uint8)
    p = str(tmp_path / f"f{ctype}.png")
    _filtered_png(p, img, ctype, [0, 1, 2, 3, 4, 4, 3, 1])
    ref = np.asarray(PIL.open(p).convert("RGBA"))
    want = {0: lambda a: np.concatenate([a.repeat(3, -1), np.full(a.shape, 255, np.uint8)], -1),
            2: lambda a: np.concatenate([a, np.full(a.shape[:2] + (1,), 255, np.uint8)], -1),
            4: lambda a: np.concatenate([a[..., :1].repeat(3, -1), a[..., 1:]], -1),
            6: lambda a: a}[ctype](img)
    assert np.array_equal(ref, want)          # the file encodes `img`
    assert np.array_equal(image.read_png(p, rgba=True), want)
    assert np.array_equal(image.read_png(p), want[..., :3])


def test_palette_with_transparency(tmp_path):
    rng = np.random.default_rng(9)
    idx = rng.integers(0, 40, (19, 27)).astype(np.uint8)
    pal = rng.integers(0, 256, (40, 3)).astype(np.uint8)
    im = PIL.fromarray(idx, "P")
    im.putpalette(pal.ravel().tolist())
    im.info["transparency"] = bytes(rng.integers(0, 256, 40).astype(np.uint8))
    p = str(tmp_path / "p.png")
    im.save(p, transparency=im.info["transparency"])
    assert np.array_equal(image.read_png(p, rgba=True), np.asarray(PIL.open(p).convert("RGBA")))


def test_pillow_encoded_files(tmp_path):
    rng = np.random.default_rng(3)
    yy, xx = np.mgrid[0:41, 0:29]
    for mode, ch in (("RGB", 3), ("RGBA", 4), ("L", 1), ("LA", 2)):
        a = ((rng.integers(0, 256, (41, 29, ch)) // 8) + (3 * xx + 5 * yy)[..., None]).astype(np.uint8)
        p = str(tmp_path / f"{mode}.png")
        PIL.fromarray(a[..., 0] if ch == 1 else a, mode).save(p, optimize=True)
        assert np.array_equal(image.read_png(p, rgba=True), np.asarray(PIL.open(p).convert("RGBA"))), mode


def test_write_read_round_trip(tmp_path):
    a = np.random.default_rng(1).integers(0, 256, (9, 14, 4)).astype(np.uint8)
    p = str(tmp_path / "w.png")
    image.write_png(p, a)
    assert np.array_equal(image.read_png(p, rgba=True), a)
    assert np.array_equal(np.asarray(PIL.open(p)), a)
