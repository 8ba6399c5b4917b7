"""Image-quality measurements of rendered frames.

rel-MSE = mean over pixels and channels of (I - I_ref)^2 / (I_ref^2 + 1e-2), reported beside the
rel-MSE of an all-black image and with the exposure normalised to a unit-mean truth.

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


def batch_report(batch_means: np.ndarray, truth: np.ndarray, truth_var: np.ndarray | None = None) -> dict:
    """rel-MSE of the plain mean (over finite pixels; non-finite ones counted) and of the
    per-pixel median of the batch means, each beside the figures that make it readable:
      * black_rel_mse: the rel-MSE of an all-black image (what a renderer that outputs nothing
        scores; with radiance far below sqrt(1e-2) it is tiny, so rel-MSE alone says little);
      * *_exposure: the same with image and truth divided by the truth's mean (unit-mean
        exposure), where the 1e-2 of the denominator is small against the signal;
      * noise_rel_mse_exposure: the exposure-normalised rel-MSE that the measured noise alone
        predicts for the plain mean: the variance of the mean of the batches (their spread
        divided by their number) plus the truth's own variance, over (truth^2 + 1e-2);
      * mean_ratio, mean_ratio_median_of_means: image mean over truth mean."""
    b = batch_means.shape[0]
    mean = batch_means.mean(0)
    finite = np.isfinite(mean).all(-1)
    mom = np.median(batch_means, axis=0)
    t = np.asarray(truth, np.float64)
    s = 1.0 / float(t.mean()) if t.mean() > 0 else 1.0
    var_mean = np.var(batch_means, axis=0, ddof=1) / b if b > 1 else np.zeros_like(t)
    tv = np.zeros_like(t) if truth_var is None else np.asarray(truth_var, np.float64)
    ok = finite[..., None] & np.isfinite(var_mean)
    noise = float(np.mean(np.where(ok, (var_mean + tv) * s * s / ((t * s) ** 2 + 1e-2), 0.0)))
    return {
        "rel_mse": rel_mse(mean[finite], t[finite]),
        "nonfinite_px": int((~finite).sum()),
        "rel_mse_median_of_means": rel_mse(mom, t),
        "black_rel_mse": rel_mse(np.zeros_like(t), t),
        "rel_mse_exposure": rel_mse(mean[finite] * s, t[finite] * s),
        "rel_mse_median_of_means_exposure": rel_mse(mom * s, t * s),
        "black_rel_mse_exposure": rel_mse(np.zeros_like(t), t * s),
        "noise_rel_mse_exposure": noise,
        "mean_ratio": float(mean[finite].mean() / t[finite].mean()) if finite.any() else float("nan"),
        "mean_ratio_median_of_means": float(mom.mean() / t.mean()),
        "batches": int(b),
    }
