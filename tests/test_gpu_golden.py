"""The HIP library against the committed golden fixtures (tests/golden/*.npz): per-pixel
radiance of every frame within 1e-4 relative, g-buffer words, reservoir flags and neighbour
sets bit for bit, in both execution modes."""
import os
import sys

import numpy as np
import pytest

from crisoptix_amd import abi
from parity_util import rel_mismatch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as G  # noqa: E402

pytestmark = pytest.mark.gpu


def load(name):
    with np.load(os.path.join(HERE, "golden", f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("exec_mode", ["wavefront", "megakernel"])
@pytest.mark.parametrize("name", list(G.CASES))
def test_hip_matches_fixture(name, exec_mode, cx_runtime, cx_ctx):
    print(f"golden {name} exec {exec_mode}", flush=True)
    g = load(name)
    case = G.CASES[name]
    sc = G.build_scene(case["scene"])
    gscene = cx_runtime.Scene(cx_ctx, sc)
    p = G.params_from_bytes(g["params"])
    W, H = case["W"], case["H"]
    ren = cx_runtime.Renderer(W, H, cx_ctx, G.camera_for(case, sc), gscene, params=p, seed=int(g["seed"]))
    ren.set_exec_mode(exec_mode)
    for f in range(case["frames"]):
        img = ren.launch_frame(G.camera_struct(g["cameras"][f]), bool(g["changed"][f])).numpy()
        bad = rel_mismatch(img, g["images"][f])
        assert not bad.any(), f"frame {f}: {int(bad.any(axis=-1).sum())} pixels differ"
    if "gbuffer" in g:
        gb = ren.gbuffer()
        assert np.array_equal(gb["primaryIntValid"], g["gbuffer"][:, 2])
        assert np.array_equal(gb["canonFlags"], g["gbuffer"][:, 3])
        st = ren.state()
        assert np.array_equal(ren.prefix(st.frontBufferIdx)["flags"], g["prefix_flags"])
    if "suffix_flags" in g:
        st = ren.state()
        assert np.array_equal(ren.suffix(st.frontBufferIdx)["flags"], g["suffix_flags"])
        gi, gd, gc = ren.neighbors()
        assert np.array_equal(gc, g["neigh_count"])
        for q in np.nonzero(gc)[0]:
            n = gc[q]
            assert np.array_equal(gi[q, :n], g["neigh_idx"][q, :n])
            assert np.array_equal(gd[q, :n].view(np.uint32), g["neigh_dist"][q, :n].view(np.uint32))
