"""Writes the golden fixtures of this directory from the oracle (oracle/), the host restatement
of the reference shader logic.  The reference ships no tests, vectors or fixtures of its own
and cannot be built here, so these fixtures record the oracle's outputs on small seeded
workloads; the oracle itself is pinned by the known-answer and closed-form tests in
tests/test_oracle_*.py.  The fixtures let the GPU tests check the HIP library against fixed
data as well as against the live oracle, and catch any later change of the oracle.

    python tests/golden/make_golden.py

Each .npz holds only plain arrays (no pickled objects): the scene recipe as a JSON string,
the launch parameters as the raw bytes of cx_launch_params, the camera of every frame, and the
outputs.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from crisoptix_amd import abi, scenes  # noqa: E402

# name -> recipe.  `scene`: scene builder and its arguments; `mode`: workload_params name;
# `over`: parameter overrides; `orbit`: rotate the camera by 2*pi/64 per frame.
CASES = {
    "pt_cornell_32x32": dict(scene=("cornell_box", dict(seed=1, tess=8)), mode="path_tracer", W=32, H=32,
                             frames=2),
    "restir_pt_cornell_32x32": dict(scene=("cornell_box", dict(seed=1, tess=8)), mode="restir_pt_cornell",
                                    W=32, H=32, frames=4),
    "cond_cornell_32x24": dict(scene=("cornell_box", dict(seed=1, tess=8)), mode="conditional_restir", W=32,
                               H=24, frames=3, over=dict(gatherRadius=0.08, suffixSpatialCount=2)),
    "cond_grid3_orbit_32x24": dict(scene=("object_grid", dict(grid=3, seed=5, ground_quads=20, ico_subdiv=2,
                                                               num_lights=4, lights="tent", trace_emitters=True)),
                                   mode="conditional_restir", W=32, H=24, frames=3, orbit=True,
                                   over=dict(gatherRadius=0.3)),
    # the Conditional ReSTIR workload reduced: 16 traced light triangles forming the tent, k = 8,
    # neighbour counts from 1 to 8 with a few hundred pixels at 8
    "cond_tent16_orbit_48x32": dict(scene=("object_grid", dict(grid=3, seed=5, ground_quads=20, ico_subdiv=2,
                                                                num_lights=16, lights="tent",
                                                                trace_emitters=True)),
                                    mode="conditional_restir", W=48, H=32, frames=3, orbit=True,
                                    over=dict(gatherRadius=0.6)),
}


def build_scene(recipe):
    name, kw = recipe
    return getattr(scenes, name)(**kw)


def camera_for(case, sc):
    W, H = case["W"], case["H"]
    if case["scene"][0] == "cornell_box":
        return scenes.cornell_camera(W, H)
    return scenes.orbit_camera(sc, W, H)


def params_for(case, size):
    p = abi.workload_params(case["mode"], case["W"], case["H"], size)
    for k, v in case.get("over", {}).items():
        setattr(p, k, v)
    return p


def params_bytes(p: abi.LaunchParams) -> np.ndarray:
    return np.frombuffer(C.string_at(C.addressof(p), C.sizeof(p)), np.uint8).copy()


def params_from_bytes(b: np.ndarray) -> abi.LaunchParams:
    p = abi.LaunchParams()
    C.memmove(C.addressof(p), np.ascontiguousarray(b, np.uint8).ctypes.data, C.sizeof(p))
    return p


def cameras(case, sc):
    """Camera data of every frame, [frames][12] floats, and the camera-changed flags."""
    cam = camera_for(case, sc)
    out, changed = [], []
    for f in range(case["frames"]):
        if case.get("orbit") and f > 0:
            cam.rotate_around_origin([0, 1, 0], 2 * np.pi / 64)
        d = cam.data()
        out.append(np.concatenate([np.asarray(d[k], np.float32) for k in ("pos", "U", "V", "W")]))
        changed.append(f == 0 or bool(case.get("orbit")))
    return np.stack(out), np.asarray(changed)


def camera_struct(row: np.ndarray) -> abi.CameraData:
    return abi.CameraData.from_dict({"pos": row[0:3], "U": row[3:6], "V": row[6:9], "W": row[9:12]})


SEED = 1234


def render_oracle(name):
    from oracle import oracle as O
    case = CASES[name]
    sc = build_scene(case["scene"])
    os_ = O.OracleScene(sc)
    p = params_for(case, os_.size)
    cams, changed = cameras(case, sc)
    ren = O.OracleRenderer(os_, case["W"], case["H"], p, seed=SEED)
    images = []
    for f in range(case["frames"]):
        images.append(ren.frame(camera_struct(cams[f]), bool(changed[f])).copy())
    out = dict(images=np.stack(images).astype(np.float32))
    if case["mode"] != "path_tracer":
        front, _, _, _ = ren.buffers()
        out["gbuffer"] = ren.gbuffer().astype(np.uint32)
        out["prefix_flags"] = np.asarray(ren.prefix(front)["flags"], np.uint32)
    if case["mode"] == "conditional_restir":
        front, _, _, _ = ren.buffers()
        out["suffix_flags"] = np.asarray(ren.suffix(front)["flags"], np.uint32)
        idx, dist, cnt = ren.neighbors()
        out["neigh_count"] = np.asarray(cnt, np.uint32)
        out["neigh_idx"] = np.asarray(idx, np.uint32)
        out["neigh_dist"] = np.asarray(dist, np.float32)
    return case, sc, p, cams, changed, out


def main():
    for name in CASES:
        case, sc, p, cams, changed, out = render_oracle(name)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, recipe=np.asarray(json.dumps(case)), params=params_bytes(p), cameras=cams,
                            changed=changed, seed=np.asarray(SEED, np.uint64), **out)
        print(f"wrote {path} ({os.path.getsize(path)} bytes), mean radiance {out['images'].mean():.5f}")


if __name__ == "__main__":
    main()
