"""The project-written sine/cosine shared by the oracle and the kernels."""
import numpy as np


def test_sincos_accuracy(orc):
    x = np.linspace(0.0, 2 * np.pi, 200001, dtype=np.float32)[:-1]
    s, c = orc.sincos(x)
    xs = x.astype(np.float64)
    assert np.max(np.abs(s - np.sin(xs))) < 3e-7
    assert np.max(np.abs(c - np.cos(xs))) < 3e-7


def test_sincos_special_values(orc):
    s, c = orc.sincos(np.asarray([0.0, np.float32(np.pi / 2), np.float32(np.pi)], np.float32))
    assert s[0] == 0.0 and c[0] == 1.0
    assert abs(s[1] - 1.0) < 1e-7 and abs(c[1]) < 1e-7
    assert abs(s[2]) < 1e-7 and abs(c[2] + 1.0) < 1e-7
    # odd / even symmetry
    x = np.random.default_rng(2).uniform(-6, 6, 1000).astype(np.float32)
    s1, c1 = orc.sincos(x)
    s2, c2 = orc.sincos(-x)
    assert np.array_equal(s1, -s2) and np.array_equal(c1, c2)
