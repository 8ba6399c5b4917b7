"""sRGB tonemap (src/kernel/tonemap.cu:7-39): the oracle's restatement against the closed-form
sRGB transfer function of IEC 61966-2-1 evaluated in double, the special values, the PNG
round trip; on the GPU, cx_tonemap against the oracle bit for bit."""
import os

import numpy as np
import pytest

from crisoptix_amd import image


def srgb_closed_form(c):
    c = np.asarray(c, np.float64)
    with np.errstate(invalid="ignore"):
        return np.where(c < 0.0031308, 12.92 * c, 1.055 * np.power(np.maximum(c, 0.0), 1.0 / 2.4) - 0.055)


def test_oracle_tonemap_matches_srgb_curve(orc):
    rng = np.random.default_rng(0)
    c = np.concatenate([np.linspace(0.0, 1.2, 200001), rng.random(100000) * 2.0,
                        [0.0031308, np.nextafter(np.float32(0.0031308), 0), 1.0, 1e-8]]).astype(np.float32)
    got = orc.tonemap(c).astype(np.int64)
    v = np.clip(srgb_closed_form(c), 0.0, 1.0) * 255.0
    exp = np.floor(v).astype(np.int64)
    # away from the truncation boundaries the float pipeline must give the exact byte
    safe = np.abs(v - np.round(v)) > 1e-4
    assert np.array_equal(got[safe], exp[safe])
    assert np.max(np.abs(got - exp)) <= 1
    # the curve is continuous at the breakpoint and monotone
    assert np.all(np.diff(orc.tonemap(np.sort(c)).astype(np.int64)) >= 0)


def test_oracle_tonemap_special_values(orc):
    c = np.asarray([-1.0, -0.0, 0.0, np.inf, -np.inf, np.nan, 1.0, 100.0], np.float32)
    got = orc.tonemap(c)
    # negative: 12.92 c < 0 clamps to 0; +inf: the lerp of 12.92 inf * 0 is NaN -> 0; NaN -> 0;
    # 1.0: 1.055f - 0.055f rounds to 1 - 2^-24 in float, so 255 x that truncates to 254 (the
    # reference's arithmetic; only values a little above 1 reach 255)
    assert list(got) == [0, 0, 0, 0, 0, 0, 254, 255]


def test_png_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    a = rng.integers(0, 256, (7, 13, 3), dtype=np.uint8)
    p = str(tmp_path / "x.png")
    image.write_png(p, a)
    assert np.array_equal(image.read_png(p), a)
    with open(p, "rb") as f:
        assert f.read(8) == b"\x89PNG\r\n\x1a\n"


@pytest.mark.gpu
def test_gpu_tonemap_bit_exact(orc, cx_runtime, cx_ctx):
    rng = np.random.default_rng(2)
    c = np.concatenate([rng.random(300000) * 1.5, rng.standard_normal(1000) * 10, np.linspace(0, 0.01, 10000),
                        [np.inf, -np.inf, np.nan, 0.0031308]]).astype(np.float32)
    d = cx_runtime.DeviceArray.from_numpy(c)
    got = cx_runtime.tonemap(cx_ctx, d)
    assert np.array_equal(got, orc.tonemap(c))
