"""The committed golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py)
against the oracle: a change of the oracle's results shows up here, and on the GPU
(test_gpu_golden.py) the HIP library is checked against the same fixed data."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as G  # noqa: E402


def load(name):
    with np.load(os.path.join(HERE, "golden", f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", list(G.CASES))
def test_fixture_recipe_is_current(name):
    import json
    g = load(name)
    assert json.loads(str(g["recipe"])) == json.loads(json.dumps(G.CASES[name]))
    case = G.CASES[name]
    sc = G.build_scene(case["scene"])
    cams, changed = G.cameras(case, sc)
    assert np.array_equal(cams.view(np.uint32), g["cameras"].view(np.uint32))
    assert np.array_equal(changed, g["changed"])


@pytest.mark.parametrize("name", list(G.CASES))
def test_oracle_reproduces_fixture(name, orc):
    g = load(name)
    case, sc, p, cams, changed, out = G.render_oracle(name)
    assert np.array_equal(G.params_bytes(p), g["params"])
    for k, v in out.items():
        assert np.array_equal(np.ascontiguousarray(v).view(np.uint8), np.ascontiguousarray(g[k]).view(np.uint8)), k
