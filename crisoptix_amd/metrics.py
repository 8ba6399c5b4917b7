"""Image-quality measurements of rendered frames.

rel-MSE = mean over pixels and channels of (I - I_ref)^2 / (I_ref^2 + 1e-2).

The reference's ReSTIR estimators are heavy-tailed: the prefix shift evaluates the reconnection
Jacobian with the source's primary-hit position in place of the vertex before the reconnection
(include/graph/restir/prefix_reuse.h:88-92) and the Jacobian is not clamped
(include/graph/restir/path_gen.h:70 is commented out), so a rare frame can put a value of 1e8 or
+inf into a pixel.  A plain running mean then measures those few frames.  `frame_batches`
therefore keeps B batch means per pixel (in float64, on the host), from which both the plain
mean and the median of the batch means (a median-of-means estimate, robust to fireflies in
fewer than half of the batches) are reported.
"""
from __future__ import annotations

import numpy as np

from . import abi


def rel_mse(img, ref, finite_only: bool = False) -> float:
    img = np.asarray(img, np.float64)
    ref = np.asarray(ref, np.float64)
    if finite_only:
        m = np.isfinite(img).all(-1)
        img, ref = img[m], ref[m]
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2)))


def frame_batches(renderer, cam: abi.CameraData, frames: int, batches: int) -> np.ndarray:
    """Renders `frames` frames of a static view (the first one resets the history) with
    accumulation off and returns the per-pixel mean of each of `batches` equal batches,
    shape [batches, H, W, 3], float64."""
    if frames % batches:
        raise ValueError("frames must be a multiple of batches")
    per = frames // batches
    p = renderer.params()
    if p.enableAccum:
        p.enableAccum = 0
        renderer.set_params(p)
    out = None
    for b in range(batches):
        acc = None
        for f in range(per):
            img = renderer.launch_frame(cam, b == 0 and f == 0).numpy().astype(np.float64)
            acc = img if acc is None else acc + img
        acc /= per
        if out is None:
            out = np.empty((batches,) + acc.shape)
        out[b] = acc
    return out


def batch_report(batch_means: np.ndarray, truth: np.ndarray) -> dict:
    """rel-MSE of the plain mean (over finite pixels; non-finite ones counted) and of the
    per-pixel median of the batch means."""
    mean = batch_means.mean(0)
    finite = np.isfinite(mean).all(-1)
    mom = np.median(batch_means, axis=0)
    return {
        "rel_mse": rel_mse(mean[finite], truth[finite]),
        "nonfinite_px": int((~finite).sum()),
        "rel_mse_median_of_means": rel_mse(mom, truth),
        "batches": int(batch_means.shape[0]),
    }
