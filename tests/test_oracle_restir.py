"""Reservoir, pairwise MIS, reconnection Jacobian and shift behaviour of the oracle
(include/graph/restir/Reservoir.h, path_gen.h, prefix_reuse.h, suffix_reuse.h)."""
import math

import numpy as np

from crisoptix_amd import abi, scenes


def test_jacobian_closed_forms(orc):
    tp = np.asarray([0, 0, 0], np.float32)
    tn = np.asarray([0, 1, 0], np.float32)
    x = np.asarray([0.3, 2.0, 0.1], np.float32)
    assert orc.jacobian(x, x, tp, tn) == 1.0
    # same direction, twice the distance: cosines equal, squared-distance ratio 1/4
    assert abs(orc.jacobian(x, 2 * x, tp, tn) - 0.25) < 1e-6
    # |cos_new / cos_old| * |x_old - y|^2 / |x_new - y|^2 with both points at distance 1
    old = np.asarray([0, 1, 0], np.float32)
    new = np.asarray([math.sin(0.5), math.cos(0.5), 0], np.float32)
    assert abs(orc.jacobian(old, new, tp, tn) - math.cos(0.5)) < 1e-6
    # degenerate: the new point on the target -> inf -> 0 (path_gen.h:69)
    assert orc.jacobian(old, tp, tp, tn) == 0.0


def test_pairwise_mis_closed_forms(orc):
    a = np.asarray([1.0, 0.5, 0.25], np.float32)
    b = np.asarray([0.2, 0.9, 0.4], np.float32)
    la = 0.2126 * 1.0 + 0.7152 * 0.5 + 0.0722 * 0.25
    lb = 0.2126 * 0.2 + 0.7152 * 0.9 + 0.0722 * 0.4
    cw, nw = orc.pairwise_mis(a, b, 1.0, 1.0, 3.0, 5.0)
    assert abs(cw - la * 3 / (la * 3 + lb * 5)) < 1e-6
    assert abs(nw - lb * 5 / (lb * 5 + la * 3)) < 1e-6
    assert abs(cw + nw - 1.0) < 1e-6
    # K neighbours: canonical term scales the neighbour side by K, the neighbour term divides by K
    cw, nw = orc.pairwise_mis(a, b, 2.0, 4.0, 1.0, 1.0)
    assert abs(cw - la / (la + lb * 2 * 4)) < 1e-6
    assert abs(nw - lb / (lb + la * 2 / 4)) < 1e-6
    # zero luminance branches (path_gen.h:23-29, 41-46)
    z = np.zeros(3, np.float32)
    assert orc.pairwise_mis(z, b, 1, 1, 1, 1) == (1.0, 1.0)
    assert orc.pairwise_mis(a, z, 1, 1, 1, 1) == (1.0, 0.0)


def test_reservoir_selection_frequencies(orc):
    w = np.asarray([1.0, 3.0, 0.0, 6.0], np.float32)
    c = np.ones(4, np.float32)
    n = 20000
    counts = np.zeros(4)
    seeds = np.random.default_rng(41).integers(0, 2**63, n)
    for s in seeds:
        idx, ws, conf = orc.reservoir_stream(w, c, int(s))
        counts[idx] += 1
    assert ws == 10.0 and conf == 4.0
    expect = w / w.sum() * n
    chi2 = ((counts - expect) ** 2 / np.maximum(expect, 1))[expect > 0].sum()
    assert counts[2] == 0
    assert chi2 < 16.3  # 3 dof, p = 0.001


def test_reservoir_nan_and_confidence_cap(orc):
    w = np.asarray([np.nan, np.inf, 2.0], np.float32)
    c = np.asarray([30.0, 30.0, 30.0], np.float32)
    idx, ws, conf = orc.reservoir_stream(w, c, 5)
    assert idx == 2 and ws == 2.0 and conf == 50.0


def _params(scene_size, mode="restir_pt_cornell"):
    p = abi.workload_params(mode, 32, 32, scene_size)
    return p


def test_prefix_identity_shift_jacobian(orc):
    """Shifting a prefix onto itself replays the same random numbers from the same primary
    hit, so the reconnection Jacobian is 1 except where the rehydrated primary vertex, whose
    interpolated normal is not re-normalised (Interaction.h:165-166), steers a replayed
    direction by a few ulps.  (The suffix shift has no such invariant: a BSDF-sampled suffix
    reconnection stores reconIdx = prefix length, path_gen.h:378, so its replay length is not
    the path's.)  The shifted contribution is not f itself either: the reference multiplies the
    reconnection vertex's BSDF twice (in postReconF, path_gen.h:246, and as brdf2,
    prefix_reuse.h:80-85)."""
    sc = scenes.cornell_box(1, 4)
    os_ = orc.OracleScene(sc)
    cam = scenes.cornell_camera(32, 32).data()
    rng = np.random.default_rng(51)
    p = _params(os_.size, "restir_pt_cornell")
    js = []
    for i in range(3000):
        d = cam["W"] + cam["U"] * rng.uniform(-1, 1) + cam["V"] * rng.uniform(-1, 1)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        out = os_.identity_shift(p, cam["pos"], d, int(rng.integers(0, 2**63)), which="prefix")
        if out[7] and out[3:6].sum() > 0:
            js.append(out[6])
    dev = np.abs(np.asarray(js) - 1.0)
    assert len(dev) > 40
    assert np.mean(dev == 0.0) > 0.9
    assert dev.max() < 0.1


def test_restir_canonical_estimator_relation(orc):
    """With reuse off, ReSTIR PT outputs f/p of one canonical prefix whose forced NEE at the last
    vertex still multiplies p by neeProb (path_gen.h:145-149).  At maxPathLen = 1 that is exactly
    the path tracer's estimator divided by neeProb, so the means agree up to that factor."""
    sc = scenes.cornell_box(1, 4)
    os_ = orc.OracleScene(sc)
    W = H = 32
    cam = abi.CameraData.from_dict(scenes.cornell_camera(W, H).data())
    pt = abi.workload_params("path_tracer", W, H, os_.size)
    pt.maxPathLen = 1
    rp = abi.workload_params("restir_pt_cornell", W, H, os_.size)
    rp.maxPathLen = 1
    rp.prefixEnableTemporal = 0
    rp.prefixEnableSpatial = 0
    means = []
    for p in (pt, rp):
        r = orc.OracleRenderer(os_, W, H, p, seed=9)
        means.append(np.mean([r.frame(cam, i == 0).mean() for i in range(48)]))
    ratio = means[1] / means[0]
    assert abs(ratio - 1 / 0.3) < 0.15 * (1 / 0.3), means


def test_unclamped_jacobian_can_make_reservoirs_infinite(orc):
    """The prefix shift evaluates the reconnection Jacobian with the source's primary-hit
    position in place of the vertex before the reconnection (prefix_reuse.h:88-92), the Jacobian
    is not clamped (path_gen.h:70 is commented out), and FinalizeGRIS divides wSum by the
    sample's target luminance (Reservoir.h:44-49) with no bound.  A primary hit and a
    reconnection vertex on the same tessellated floor make cos_old rounding noise, the RIS weight
    finite but huge, and the final division overflows to +inf: the reservoir then stores
    wSum = +inf beside a finite, tiny f, and the unguarded ReSTIR PT output wSum * f is +inf.
    This is the reference's estimator, so the GPU frames reproduce it bit for bit."""
    sc = scenes.cornell_box(1, 8)
    os_ = orc.OracleScene(sc)
    W = H = 64
    cd = abi.CameraData.from_dict(scenes.cornell_camera(W, H).data())
    p = abi.workload_params("restir_pt_cornell", W, H, os_.size)
    p.maxPathLen = 4
    r = orc.OracleRenderer(os_, W, H, p, seed=99)
    for i in range(64):
        out = r.frame(cd, i == 0)
        bad = ~np.isfinite(out).all(-1)
        if bad.any():
            break
    assert bad.any(), "expected the unclamped Jacobian to overflow a reservoir within 64 frames"
    assert bad.sum() <= 4
    front = r.prefix(r.buffers()[0])
    for px in np.flatnonzero(bad.ravel()):
        rec = front[px]
        assert np.isposinf(rec["wSum"])
        assert np.isfinite(rec["f"]).all() and 0 < float(np.max(rec["f"])) < 1e-3
