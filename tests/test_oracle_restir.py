"""Reservoir, pairwise MIS, reconnection Jacobian and shift behaviour of the oracle
(include/graph/restir/Reservoir.h, path_gen.h, prefix_reuse.h, suffix_reuse.h)."""
import math

import numpy as np
import pytest

from crisoptix_amd import abi, scenes
from parity_util import tent_grid_scene


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


def test_pairwise_mis_partition_of_unity(orc):
    """For a path x of the canonical domain and y = T^-1(x) of a neighbour's, the canonical weight
    of x in that pair (ComputeCanonicalPairwiseMISWeight with J_{x->y}) and the weight x gets as
    the neighbour's shifted sample (ComputeNeighborPairwiseMISWeight with J_{y->x} = 1 / J_{x->y})
    sum to one for any target values, confidences, K and Jacobian (path_gen.h:15-48): the pairwise
    MIS of the spatial passes and the final gather is a partition of unity pair by pair.  A
    Jacobian applied in the wrong direction, K on the wrong side or swapped confidences breaks it."""
    rng = np.random.default_rng(17)
    worst = 0.0
    for _ in range(2000):
        fc = rng.uniform(0.01, 5.0, 3).astype(np.float32)
        fn = rng.uniform(0.01, 5.0, 3).astype(np.float32)
        J = float(np.exp(rng.normal(0.0, 1.5)))
        K = float(rng.integers(1, 9))
        cM = float(rng.uniform(0.5, 3.0))
        nM = float(rng.uniform(1.0, 50.0))
        cw, _ = orc.pairwise_mis(fc, fn, J, K, cM, nM)
        _, nw = orc.pairwise_mis(fc, fn, np.float32(1.0 / J), K, cM, nM)
        worst = max(worst, abs(cw + nw - 1.0))
    assert worst < 1e-5, worst
    # the same two calls with the Jacobian not inverted are not a partition
    cw, _ = orc.pairwise_mis(np.ones(3, np.float32), np.ones(3, np.float32), 4.0, 1.0, 1.0, 1.0)
    _, nw = orc.pairwise_mis(np.ones(3, np.float32), np.ones(3, np.float32), 4.0, 1.0, 1.0, 1.0)
    assert abs(cw + nw - 1.0) > 0.5


@pytest.mark.parametrize("max_len", [3, 4])
def test_suffix_identity_shift_at_nee_reconnection(orc, max_len):
    """A suffix that took NEE at its first vertex (the prefix's last vertex) is reconnected there:
    reconIdx = 1 = its length (path_gen.h:320-341).  Shifting it onto its own last prefix vertex
    replays no vertex (suffix_reuse.h:34), so the reconnection Jacobian compares a vertex with
    itself and is exactly 1 (path_gen.h:59-72), and the shifted value brdf(to the reconnection
    point) * colour / d^2 is the suffix's f up to the rounding of the reconnection hit point
    (suffix_reuse.h:63-104).  Light traced as geometry (C9'), so the reconnection ray ends on the
    light as in the reference's scenes."""
    sc = scenes.cornell_box(1, 4, trace_emitters=True)
    os_ = orc.OracleScene(sc)
    cam = scenes.cornell_camera(32, 32).data()
    rng = np.random.default_rng(53 + max_len)
    p = abi.workload_params("conditional_restir", 32, 32, os_.size)
    p.maxPathLen = max_len
    n = 0
    for i in range(4000):
        d = cam["W"] + cam["U"] * rng.uniform(-1, 1) + cam["V"] * rng.uniform(-1, 1)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        out = os_.identity_shift(p, cam["pos"], d, int(rng.integers(0, 2**63)), which="suffix")
        if not out[7] or out[8] != 1 or out[9] != 1:
            continue
        f, fs, J = out[0:3].astype(np.float64), out[3:6].astype(np.float64), out[6]
        if f.sum() <= 0:
            continue
        n += 1
        assert J == 1.0
        assert np.allclose(fs, f, rtol=1e-4, atol=0.0), (f, fs)
    assert n > 300


@pytest.mark.parametrize("scene", ["cornell_tl", "grid"])
def test_suffix_identity_shift_at_bsdf_reconnection(orc, scene):
    """With prefix length 1, a suffix whose first traced vertex is rough and far enough to be its
    reconnection vertex (path_gen.h:362-381) stores reconIdx = 1, the prefix length (:378), so
    shifting it onto its own last prefix vertex replays no vertex (suffix_reuse.h:34) and
    reconnects from that vertex straight to the reconnection vertex: the Jacobian compares the
    last prefix vertex with itself and is exactly 1 (path_gen.h:59-72), and the shifted value
    brdf(last prefix vertex -> reconnection vertex) * brdf(reconnection vertex -> reconOutDir) *
    postReconF (suffix_reuse.h:63-104) is the suffix's f up to the rounding of the vertex
    positions (the sampled directions there and the reconnection direction differ in the last
    bits).  A wrong replay count, a BSDF evaluated with the wrong directions or postReconF taking
    the reconnection vertex's own BSDF would each fail it.  The shift's visibility ray to the
    reconnection vertex ends on that vertex's own surface at about t = tmax, so it is occluded in
    part of the cases, as in the reference (C5 counts a hit at tmax), and the shift then returns
    zero; the identity is checked on the others.  On the Cornell box (flat faces) f comes back
    within 1e-4; on the object grid the reconnection vertex is rehydrated with its interpolated
    normal left unnormalised (Interaction.h:165-169, C7) while the suffix was traced with the
    normalised one (mesh_hit.cu:62-66), so its BSDF differs by up to ~10 % there and only J = 1 is
    exact."""
    sc = scenes.cornell_box(1, 4, trace_emitters=True) if scene == "cornell_tl" else tent_grid_scene()
    os_ = orc.OracleScene(sc)
    cam = (scenes.cornell_camera(32, 32) if scene == "cornell_tl" else scenes.orbit_camera(sc, 32, 32)).data()
    rng = np.random.default_rng(61)
    p = abi.workload_params("conditional_restir", 32, 32, os_.size)
    p.prefixLen = 1
    p.maxPathLen = 5
    n = occluded = 0
    for i in range(6000):
        d = cam["W"] + cam["U"] * rng.uniform(-1, 1) + cam["V"] * rng.uniform(-1, 1)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        out = os_.identity_shift(p, cam["pos"], d, int(rng.integers(0, 2**63)), which="suffix")
        if not out[7] or out[11] != 1 or out[10] != 1:
            continue
        assert out[8] == 1   # reconIdx = the prefix length
        f, fs, J = out[0:3].astype(np.float64), out[3:6].astype(np.float64), out[6]
        if f.sum() <= 0:
            continue
        if J == 0.0 and not fs.any():   # the visibility ray met the reconnection surface at tmax
            occluded += 1
            continue
        n += 1
        assert J == 1.0
        if scene == "cornell_tl":
            assert np.allclose(fs, f, rtol=1e-4, atol=0.0), (f, fs)
    assert n > 100 and occluded < 4 * n, (n, occluded)


def test_gather_combination_limits(orc):
    """The final gather's combination (final_gather.cu:135-187), fed injected neighbour terms:
    * neighbours that cannot produce the canonical path and that the canonical suffix cannot reach
      (both shifted values zero) leave the canonical estimate unchanged: each pair's canonical
      weight is 1, so canonMisWeight = K + 1 cancels the final 1/(K + 1), for K = 1..8;
    * a zero neighbour weight removes that neighbour's own term but keeps its pairwise share of the
      canonical weight;
    * one neighbour identical to the canonical suffix (shift J = 1, confidence 1, W = 1/p) gives
      (1/4 + 3/2)/2 = 7/8 of the canonical estimate: the neighbour term is divided by N + 1 = 2 at
      line 181 and again by K + 1 = 2 at line 184, where the pair's weights (1/2 each) alone would
      return the canonical estimate; the oracle keeps the reference's double division."""
    pre_f, pre_p = np.asarray([0.6, 0.5, 0.4], np.float32), 0.7
    suf_f, suf_p = np.asarray([0.3, 0.2, 0.25], np.float32), 0.9
    canon = (pre_f * suf_f / np.float32(pre_p * suf_p)).astype(np.float32)
    rng = np.random.default_rng(3)
    for K in range(1, 9):
        t = np.zeros((K, 13), np.float32)
        t[:, 3] = rng.uniform(0.5, 2.0, K)        # Jacobians
        t[:, 4:7] = rng.uniform(0.1, 1.0, (K, 3))  # neighbours' own values
        t[:, 7] = rng.uniform(0.5, 3.0, K)        # reservoir weights
        t[:, 8] = rng.uniform(1.0, 50.0, K)       # confidences
        t[:, 12] = rng.uniform(0.5, 2.0, K)
        r = orc.gather_combine(pre_f, pre_p, suf_f, canon, t)
        assert np.allclose(r, canon, rtol=2e-7, atol=0), (K, r, canon)
    # a neighbour with zero reservoir weight next to one that reaches the canonical path
    t = np.zeros((2, 13), np.float32)
    t[:, 0:3] = suf_f
    t[:, 3] = 1.0
    t[:, 4:7] = suf_f
    t[:, 8] = 1.0
    t[:, 9:12] = suf_f
    t[:, 12] = 1.0
    t[0, 7] = 1.0 / suf_p
    with_zero = orc.gather_combine(pre_f, pre_p, suf_f, canon, t)
    t[1, 7] = 1.0 / suf_p
    both = orc.gather_combine(pre_f, pre_p, suf_f, canon, t)
    assert np.all(both > with_zero)
    # one neighbour identical to the canonical suffix
    r1 = orc.gather_combine(pre_f, pre_p, suf_f, canon, t[:1])
    assert np.allclose(r1, canon * 0.875, rtol=1e-6), (r1 / canon)
