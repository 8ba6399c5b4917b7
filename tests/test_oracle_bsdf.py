"""GGX BSDF of shader/brdf.cu pinned by closed forms and Monte-Carlo identities."""
import math

import numpy as np
import pytest

from crisoptix_amd import scenes


def plate(albedo, rough):
    """One triangle in the z = 0 plane, normal +z, one material."""
    b = scenes._Builder()
    m = b.add_material(albedo, rough)
    b.add_mesh(np.asarray([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float64),
               np.asarray([[0, 0, 1]] * 3, np.float64), np.asarray([[0, 1, 2]]), m)
    return b.build([([[(0, 0, 5), (1, 0, 5), (0, 1, 5)]], (1.0, 1.0, 1.0))])


def inter_rows(n, in_dir):
    rows = np.zeros((n, 8), np.float32)
    rows[:, 0] = 0  # primitive 0
    rows[:, 1] = 1 / 3
    rows[:, 2] = 1 / 3
    rows[:, 3:6] = in_dir
    rows[:, 6] = 0  # traced (normalised) normal
    return rows


@pytest.mark.parametrize("rough", [0.3, 1.0])
def test_normal_incidence_closed_form(orc, rough):
    alb = np.asarray([0.6, 0.4, 0.2])
    sc = orc.OracleScene(plate(alb, rough))
    out = sc.ggx_eval(inter_rows(1, [0, 0, -1]), np.asarray([[0, 0, 1]], np.float32))[0]
    a2 = rough * rough
    D = 1.0 / (math.pi * a2)
    V = 0.25
    # specF0 = mix(diffColor, specF0, metal = 0) = albedo (brdf.cu:137-138)
    brdf = alb / math.pi + D * V * alb
    diffProb = 0.5
    pdf = diffProb / math.pi + (1 - diffProb) * D / 4.0
    assert np.allclose(out[0:3], brdf, rtol=2e-6)
    assert abs(out[3] - pdf) < 2e-6 * pdf
    assert out[7] == np.float32(rough)


def test_back_facing_is_zero(orc):
    sc = orc.OracleScene(plate([0.5, 0.5, 0.5], 0.5))
    out = sc.ggx_eval(inter_rows(1, [0, 0, 1]), np.asarray([[0, 0, 1]], np.float32))[0]
    assert (out[0:4] == 0).all()
    smp, _ = sc.ggx_sample(inter_rows(1, [0, 0, 1]), np.asarray([123], np.uint64))
    assert (smp[0, 0:7] == 0).all()


@pytest.mark.parametrize("rough", [0.2, 0.7])
def test_sample_consistent_with_eval(orc, rough):
    sc = orc.OracleScene(plate([0.7, 0.5, 0.3], rough))
    n = 4000
    v = np.asarray([0.3, -0.2, -0.9])
    v /= np.linalg.norm(v)
    rows = inter_rows(n, v)
    states = np.random.default_rng(3).integers(0, 2**63, n).astype(np.uint64)
    smp, _ = sc.ggx_sample(rows, states)
    ok = smp[:, 6] > 0
    ev = sc.ggx_eval(rows[ok], smp[ok, 0:3])
    assert np.allclose(ev[:, 0:3], smp[ok, 3:6], rtol=1e-5, atol=1e-7)
    assert np.allclose(ev[:, 3], smp[ok, 6], rtol=1e-5)


@pytest.mark.parametrize("rough", [0.25, 1.0])
def test_sampling_histogram_matches_pdf(orc, rough):
    """Directions drawn by ggx_sample follow the returned pdf: the histogram of cos(theta)
    agrees with the integral of the eval pdf over each band (uniform-hemisphere quadrature)."""
    sc = orc.OracleScene(plate([0.5, 0.5, 0.5], rough))
    v = np.asarray([0.0, 0.6, -0.8])
    n = 200000
    rows = inter_rows(n, v)
    smp, _ = sc.ggx_sample(rows, np.random.default_rng(4).integers(0, 2**63, n).astype(np.uint64))
    ok = smp[:, 6] > 0
    bins = np.linspace(0, 1, 11)
    hist = np.histogram(smp[ok, 2], bins)[0] / n
    m = 400000
    rng = np.random.default_rng(5)
    z = rng.random(m)
    phi = rng.random(m) * 2 * np.pi
    r = np.sqrt(1 - z * z)
    dirs = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1).astype(np.float32)
    pdf = sc.ggx_eval(inter_rows(m, v), dirs)[:, 3].astype(np.float64)
    expect = np.asarray([(pdf * ((z >= lo) & (z < hi))).mean() * 2 * np.pi for lo, hi in zip(bins[:-1], bins[1:])])
    assert np.allclose(hist, expect, atol=4e-3, rtol=0.04), (hist, expect)


def test_reciprocity(orc):
    sc = orc.OracleScene(plate([0.6, 0.6, 0.6], 0.4))
    rng = np.random.default_rng(6)
    for _ in range(50):
        a = rng.normal(size=3)
        b = rng.normal(size=3)
        a[2], b[2] = abs(a[2]) + 0.1, abs(b[2]) + 0.1
        a /= np.linalg.norm(a)
        b /= np.linalg.norm(b)
        f_ab = sc.ggx_eval(inter_rows(1, -a), b[None].astype(np.float32))[0]
        f_ba = sc.ggx_eval(inter_rows(1, -b), a[None].astype(np.float32))[0]
        assert np.allclose(f_ab[0:3] / b[2], f_ba[0:3] / a[2], rtol=2e-5)
