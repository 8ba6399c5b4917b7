// Path-replay adjoint of the frame estimator (ReSTIR PT and the path tracer).
//
// The reference has no differentiable rendering; this is the adjoint of the estimator the frame
// loop already computes, with respect to per-material albedo and roughness and per-emitter
// radiance.  For ReSTIR PT the pixel estimate is I = W * f(x), W the unbiased contribution
// weight of the reservoir-selected path x (the final reservoir's wSum after FinalizeGRIS) and f
// its contribution in the pixel's domain.  The gradient estimate is W * df(x)/dtheta with W and
// the path x held fixed (W is a valid contribution weight for any integrand over path space, so
// this is an unbiased estimate of the gradient of the integral of f when sampling is detached
// from theta, as in path-replay backpropagation).  The path tracer's estimate is
// f(x) / p(x) with p detached.
//
// No tape is stored: each pixel re-traces its selected path from the reservoir record.  A path
// selected as the pixel's own canonical prefix is replayed from its primary hit with the
// record's PCG32 state.  A path shifted in from another pixel or frame is replayed as the hybrid
// shift that produced it (prefix_reuse.h:10-99: random replay from the pixel's primary hit up to
// the reconnection vertex, then the reconnection), and its post-reconnection part, which the
// record only holds as the product postReconF, is re-traced from the reconnection vertex with
// the source's PCG32 advanced past the draws of the vertices before it (four per vertex: the NEE
// decision and three for the BSDF sample), exactly as TracePrefix drew them (path_gen.h:74-250).
//
// Directions, hits and every discrete decision come from the scene's sampling material table
// (the parameters the frame was rendered with); BSDF values, emitter colours and their
// derivatives come from the evaluation table.  With the two tables equal, the replayed W * f is
// bit-identical to the frame's output; with the evaluation table perturbed, the replay is the
// frozen-path primal whose central differences the adjoint must match.
#include "cx_internal.h"
#include "cx_restir_dev.h"

namespace cx {

static constexpr int ABLOCK = 128;
static constexpr int MAX_FACTORS = 16;
static constexpr uint32_t NO_EMITTER = 0xFFFFFFFFu;

// One multiplicative factor of f: its value, d/d(albedo of `mat`) per channel, d/d(roughness
// of `mat`) per channel and, for an NEE term, d/d(colour of `emit`) per channel.
struct Factor {
    V3 v, dA, dR, dC;
    uint32_t mat, emit;
};

struct Factors {
    Factor f[MAX_FACTORS];
    int n;
    bool overflow;
    __device__ __forceinline__ void push(const Factor& x) {
        if (n < MAX_FACTORS) f[n++] = x;
        else overflow = true;
    }
};

__device__ __forceinline__ Mat load_mat_from(const DMaterial* mats, const DScene& S, const Inter& it) {
    return resolve_mat(S, mats[S.triMat[it.prim]], it.prim, it.bu, it.bv);
}

// BSDF value (the expression of ggx_eval / ggx_sample, brdf.cu:115-338) and its derivatives.
// With F0 = a (1 - m) + s m (metal m; m = 0 unless the material maps OccRoughMetal):
// value_c = (a_c / pi + D V (F0_c + (1 - F0_c) P)) nL, so
// d/da_c = (1/pi + D V (1 - P) (1 - m)) nL and d/dr = (dD/dr V + D dV/dr) F_c nL, with
// D = a2 / (pi q^2), q = nH^2 (a2 - 1) + 1, V = 0.5 / (nL s_v + nV s_l),
// s_v = sqrt(nV^2 (1 - a2) + a2), s_l = sqrt(nL^2 (1 - a2) + a2), a2 = r^2.  A diffuse colour
// or roughness read from a texture is not the material's parameter: its derivative is zero.
struct BrdfD {
    V3 val, dA, dR;
};

__device__ __forceinline__ BrdfD brdf_d(const Mat& m, V3 normal, V3 viewDir, V3 lightDir) {
    const float nDotV = dot(normal, viewDir);
    const float nDotL = dot(normal, lightDir);
    const V3 diffBrdf = m.diff / PI_F;
    const V3 halfway = normalize(lightDir + viewDir);
    const float nDotH = dot(halfway, normal);
    const float lDotH = dot(halfway, lightDir);
    const V3 specBrdf = spec_term(m, nDotH, nDotL, nDotV, lDotH);
    BrdfD r;
    r.val = (diffBrdf + specBrdf) * nDotL;
    float dA = 1.0f / PI_F;
    V3 dR = v3(0.0f);
    if ((double)dot(m.spec, m.spec) > 1e-6) {
        const float a2 = m.rough * m.rough;
        const float d = d_ggx(nDotH, m.rough);
        const float v = v_smith(nDotL, nDotV, m.rough);
        const float p = pow5(gmax(0.0f, 1.0f - lDotH));
        dA += d * v * (1.0f - p) * (1.0f - m.metal);
        const float q = nDotH * nDotH * (a2 - 1.0f) + 1.0f;
        const float dDda2 = (q - 2.0f * a2 * nDotH * nDotH) / (PI_F * q * q * q);
        const float sv = sqrtf(nDotV * nDotV * (1.0f - a2) + a2);
        const float sl = sqrtf(nDotL * nDotL * (1.0f - a2) + a2);
        const float lam = nDotL * sv + nDotV * sl;
        const float dlam = nDotL * (1.0f - nDotV * nDotV) / (2.0f * sv) + nDotV * (1.0f - nDotL * nDotL) / (2.0f * sl);
        const float dVda2 = -0.5f * dlam / (lam * lam);
        const float dDVdr = 2.0f * m.rough * (dDda2 * v + d * dVda2);
        const V3 F = m.spec + (v3(1.0f) - m.spec) * p;
        dR = F * dDVdr;
    }
    r.dA = m.diffParam ? v3(dA * nDotL) : v3(0.0f);
    r.dR = m.roughParam ? dR * nDotL : v3(0.0f);
    return r;
}

// sample_emitter (sample_emitter.h:3-12) that also reports the emitter index.
__device__ __forceinline__ V3 sample_emitter_idx(const DScene& S, Rng& rng, uint32_t& ei) {
    const uint64_t e = rng.u64() % (uint64_t)S.numEmitters;
    const uint32_t f0 = S.emitOff[e], f1 = S.emitOff[e + 1];
    const uint64_t face = f0 + rng.u64() % (uint64_t)(f1 - f0);
    const float* fv = S.emitFaces + 9 * face;
    const V3 v0 = V3{fv[0], fv[1], fv[2]}, v1 = V3{fv[3], fv[4], fv[5]}, v2 = V3{fv[6], fv[7], fv[8]};
    float r0 = rng.f();
    float r1 = rng.f();
    float r2 = rng.f();
    const float rSum = r0 + r1 + r2;
    r0 /= rSum;
    r1 /= rSum;
    r2 /= rSum;
    ei = (uint32_t)e;
    return r0 * v0 + r1 * v1 + r2 * v2;
}

struct AdjCtx {
    const FrameArgs& A;
    const AdjArgs& G;
    TraceCtx& tc;
    double* lds;   // the block's gradient sums (G.ldsSlots doubles), or null
};

// NEE at `it` toward a sampled emitter point (path_gen.h:144-211 and :420-437).  The factor is
// (brdf * colour) * distFactor; the raw parts are returned for callers that combine them in
// another order.  `emis` is the material emission ggx_eval reports (zero in its zero branch).
// False if the replay leaves the recorded path (occluded, or a zero BSDF where the path had a
// positive one).
struct NeeOut {
    Factor fac;
    V3 brdf, color, emis;
    float distFactor;
};

__device__ __forceinline__ bool nee_factor(const AdjCtx& C, const Inter& it, Rng& rng, bool requireValue, NeeOut& o) {
    const DScene& S = C.A.S;
    uint32_t ei = 0;
    const V3 lp = sample_emitter_idx(S, rng, ei);
    const V3 lightDir = normalize(lp - it.pos);
    const float dist = length(lp - it.pos);
    if (trace_occluded(S, C.tc, it.pos, lightDir, 1e-3f, dist)) return false;
    o.color = V3{C.G.evalEmit[3 * ei], C.G.evalEmit[3 * ei + 1], C.G.evalEmit[3 * ei + 2]};
    const Mat m = load_mat_from(C.G.evalMats, S, it);
    o.distFactor = 1.0f / (dist * dist);
    const float nDotV = dot(it.nrm, -it.dir), nDotL = dot(it.nrm, lightDir);
    o.fac.mat = S.triMat[it.prim];
    o.fac.emit = ei;
    if (nDotV <= 0.0f || nDotL < 0.0f) {   // ggx_eval's zero branch
        if (requireValue) return false;
        o.brdf = o.emis = v3(0.0f);
        o.fac.v = o.fac.dA = o.fac.dR = o.fac.dC = v3(0.0f);
        return true;
    }
    const BrdfD b = brdf_d(m, it.nrm, -it.dir, lightDir);
    o.brdf = b.val;
    o.emis = m.emis;
    o.fac.v = b.val * o.color * o.distFactor;
    o.fac.dA = b.dA * o.color * o.distFactor;
    o.fac.dR = b.dR * o.color * o.distFactor;
    o.fac.dC = b.val * o.distFactor;
    return true;
}

// TracePrefix in ReSTIR PT form (every vertex draws the NEE decision, the last one forces
// NEE) replayed from vertex traceIdx0.  If haveIt, `it` is that vertex; otherwise it is traced
// from (pos, dir).  `rng` stands at that vertex's NEE decision.  Appends the factors and
// multiplies them onto `prod` one by one, in the forward's order (so the product rounds as
// TracePrefix's f and postReconF did); `emis` receives the emission term TracePrefix adds at a
// first-vertex NEE (path_gen.h:204).
__device__ __forceinline__ bool replay_prefix_tail(const AdjCtx& C, Inter it, bool haveIt, V3 pos, V3 dir,
                                                   uint32_t traceIdx0, Rng& rng, Factors& F, V3& prod, V3& emis,
                                                   bool& neeAtFirst) {
    const DScene& S = C.A.S;
    const cx_launch_params& lp = C.A.lp;
    const uint32_t maxLen = (uint32_t)lp.maxPathLen;
    emis = v3(0.0f);
    neeAtFirst = false;
    for (uint32_t traceIdx = traceIdx0; traceIdx < maxLen; ++traceIdx) {
        if (!(haveIt && traceIdx == traceIdx0)) {
            trace_surface(S, C.tc, pos, dir, 1e-3f, 1e16f, it);
            if (!it.valid) return false;
        }
        const bool forceNee = traceIdx == maxLen - 1;
        if (rng.f() < lp.neeProb || forceNee) {
            NeeOut no;
            if (!nee_factor(C, it, rng, true, no)) return false;
            prod = prod * no.fac.v;
            F.push(no.fac);
            if (traceIdx == 0) {
                emis = no.emis * no.distFactor;
                neeAtFirst = true;
            }
            return true;
        }
        const BrdfSample bs = ggx_sample(S, it, rng);
        if (bs.pdf <= 0.0f) return false;
        const BrdfD b = brdf_d(load_mat_from(C.G.evalMats, S, it), it.nrm, -it.dir, bs.outDir);
        prod = prod * b.val;
        Factor bf_;
        bf_.v = b.val;
        bf_.dA = b.dA;
        bf_.dR = b.dR;
        bf_.dC = v3(0.0f);
        bf_.mat = S.triMat[it.prim];
        bf_.emit = NO_EMITTER;
        F.push(bf_);
        pos = it.pos;
        dir = bs.outDir;
    }
    return false;   // a valid ReSTIR PT prefix always ends in NEE
}

// Interaction of a packed hit as trace_surface produced it (normalised normal).
__device__ __forceinline__ Inter traced_inter(const DScene& S, const Packed& pk) {
    Inter r = rehydrate(S, pk);
    if (r.valid) r.nrm = normalize(r.nrm);
    return r;
}

// Post-reconnection part of a source prefix (the factors whose product is its postReconF).
__device__ __forceinline__ bool replay_post_recon(const AdjCtx& C, const Prefix& src, Factors& F, V3& prc) {
    const DScene& S = C.A.S;
    const uint32_t r = src.recon_idx(), L = src.length();
    if (r < 2 || L < r || L > 32) return false;
    Rng rng;
    rng.s = src.rng;
    for (uint32_t i = 0; i < 4u * (r - 1u); i++) rng.u32();
    const Inter y = traced_inter(S, src.recon);
    if (!y.valid) return false;
    V3 emis;
    bool first;
    if (L == r) {   // reconnection at the NEE vertex (path_gen.h:165-187)
        rng.f();
        NeeOut no;
        if (!nee_factor(C, y, rng, true, no)) return false;
        prc = v3(1.0f) * no.fac.v;
        F.push(no.fac);
        return true;
    }
    rng.f();
    const BrdfSample bs = ggx_sample(S, y, rng);
    if (bs.pdf <= 0.0f) return false;
    const BrdfD b = brdf_d(load_mat_from(C.G.evalMats, S, y), y.nrm, -y.dir, bs.outDir);
    Factor yf;
    yf.v = b.val;
    yf.dA = b.dA;
    yf.dR = b.dR;
    yf.dC = v3(0.0f);
    yf.mat = S.triMat[y.prim];
    yf.emit = NO_EMITTER;
    F.push(yf);
    prc = v3(1.0f) * b.val;
    Inter none;
    none.valid = false;
    return replay_prefix_tail(C, none, false, y.pos, bs.outDir, r, rng, F, prc, emis, first);
}

// The hybrid shift of `src` onto the pixel's primary hit (prefix_reuse.h:10-99) with values
// from the evaluation table; f = max(0, throughput * postReconF).
__device__ __forceinline__ bool replay_shift(const AdjCtx& C, const Prefix& src, Factors& F, V3& f, V3& mask) {
    const DScene& S = C.A.S;
    if (!src.valid() || src.recon_idx() < 2) return false;
    Inter curr = rehydrate(S, src.primary);
    Inter srcRecon = rehydrate(S, src.recon);
    if (!curr.valid || !srcRecon.valid) return false;
    const int rvc = (int)src.recon_idx() - 2;
    Rng srng;
    srng.s = src.rng;
    V3 thr = v3(1.0f);
    for (int i = 0; i < rvc; ++i) {
        srng.f();
        const BrdfSample bs = ggx_sample(S, curr, srng);
        if (bs.pdf <= 0.0f) return false;
        const BrdfD b = brdf_d(load_mat_from(C.G.evalMats, S, curr), curr.nrm, -curr.dir, bs.outDir);
        thr = thr * b.val;
        F.push(Factor{b.val, b.dA, b.dR, v3(0.0f), S.triMat[curr.prim], NO_EMITTER});
        trace_surface(S, C.tc, curr.pos, bs.outDir, 1e-3f, 1e16f, curr);
        if (!curr.valid) return false;
    }
    const V3 reconVec = srcRecon.pos - curr.pos;
    const float reconLen = length(reconVec);
    const V3 reconDir = normalize(reconVec);
    if (isinf(reconLen) || any_bad(reconDir)) return false;
    if (trace_occluded(S, C.tc, curr.pos, reconDir, 1e-3f, reconLen)) return false;
    if (dot(curr.nrm, -curr.dir) <= 0.0f || dot(curr.nrm, reconDir) < 0.0f) return false;
    const BrdfD b1 = brdf_d(load_mat_from(C.G.evalMats, S, curr), curr.nrm, -curr.dir, reconDir);
    thr = thr * b1.val;
    F.push(Factor{b1.val, b1.dA, b1.dR, v3(0.0f), S.triMat[curr.prim], NO_EMITTER});
    srcRecon.dir = reconDir;
    if (dot(srcRecon.nrm, -srcRecon.dir) <= 0.0f || dot(srcRecon.nrm, src.reconOutDir) < 0.0f) return false;
    const BrdfD b2 = brdf_d(load_mat_from(C.G.evalMats, S, srcRecon), srcRecon.nrm, -srcRecon.dir, src.reconOutDir);
    thr = thr * b2.val;
    F.push(Factor{b2.val, b2.dA, b2.dR, v3(0.0f), S.triMat[srcRecon.prim], NO_EMITTER});
    V3 prc;
    if (!replay_post_recon(C, src, F, prc)) return false;
    const V3 x = thr * prc;
    f = V3{gmax(0.0f, x.x), gmax(0.0f, x.y), gmax(0.0f, x.z)};
    mask = V3{x.x > 0.0f ? 1.0f : 0.0f, x.y > 0.0f ? 1.0f : 0.0f, x.z > 0.0f ? 1.0f : 0.0f};
    return !F.overflow;
}

// Adds v to gradient slot `slot` (layout of AdjArgs::ldsSlots) of the block's LDS sums, or to
// the global element g when the block has none.
__device__ __forceinline__ void grad_add(double* lds, uint32_t slot, double* g, double v) {
    if (lds) atomicAdd(&lds[slot], v);
    else atomicAdd(g, v);
}

// d(u . prod_k F_k) / dtheta scattered to the gradient arrays: each factor's derivative times
// the product of all the others (prefix and suffix products, no division).
__device__ __forceinline__ void contract(const Factors& F, V3 u, const AdjArgs& G, double* lds) {
    V3 suf[MAX_FACTORS + 1];
    suf[F.n] = v3(1.0f);
    for (int k = F.n - 1; k >= 0; --k) suf[k] = F.f[k].v * suf[k + 1];
    V3 pre = v3(1.0f);
    for (int k = 0; k < F.n; ++k) {
        const Factor& x = F.f[k];
        const V3 w = u * (pre * suf[k + 1]);
        const uint32_t M = G.numMaterials;
        if (G.gAlbedo) {
            const V3 ga = w * x.dA;
            grad_add(lds, 3 * x.mat + 0, &G.gAlbedo[3 * (size_t)x.mat + 0], (double)ga.x);
            grad_add(lds, 3 * x.mat + 1, &G.gAlbedo[3 * (size_t)x.mat + 1], (double)ga.y);
            grad_add(lds, 3 * x.mat + 2, &G.gAlbedo[3 * (size_t)x.mat + 2], (double)ga.z);
        }
        if (G.gRough) grad_add(lds, 3 * M + x.mat, &G.gRough[x.mat], (double)dot(w, x.dR));
        if (G.gEmit && x.emit != NO_EMITTER) {
            const V3 ge = w * x.dC;
            grad_add(lds, 4 * M + 3 * x.emit + 0, &G.gEmit[3 * (size_t)x.emit + 0], (double)ge.x);
            grad_add(lds, 4 * M + 3 * x.emit + 1, &G.gEmit[3 * (size_t)x.emit + 1], (double)ge.y);
            grad_add(lds, 4 * M + 3 * x.emit + 2, &G.gEmit[3 * (size_t)x.emit + 2], (double)ge.z);
        }
        pre = pre * x.v;
    }
}

__device__ __forceinline__ void status_add(const AdjArgs& G, int i) {
    if (G.status) atomicAdd(&G.status[i], 1u);
}

// One pixel of ReSTIR PT: the final reservoir of the frame (front buffer).
__device__ __forceinline__ void adjoint_restir_pt(const AdjCtx& C, uint32_t p, V3 g) {
    const FrameArgs& A = C.A;
    const AdjArgs& G = C.G;
    V3 out = v3(0.0f);
    const PrefixRes res = load_prefix(A.prefixRes[A.front], A.P, p);
    const bool primaryValid = A.gbuf[p].z != 0u;
    if (primaryValid && res.s.valid() && res.wSum != 0.0f) {
        Rng canon = rng_seed(tea64(p, frame_random(A)));
        canon.f();
        canon.f();
        Factors F;
        F.n = 0;
        F.overflow = false;
        V3 f = v3(0.0f), mask = v3(1.0f);
        bool ok;
        if (res.s.rng == canon.s) {   // the pixel's own canonical prefix of this frame
            Rng rng;
            rng.s = res.s.rng;
            V3 prod = v3(1.0f), emis = v3(0.0f);
            bool neeFirst = false;
            const Inter prim = traced_inter(A.S, res.s.primary);
            ok = prim.valid && replay_prefix_tail(C, prim, true, v3(0.0f), v3(0.0f), 0u, rng, F, prod, emis, neeFirst) &&
                 !F.overflow;
            f = (neeFirst && res.s.length() == 1) ? prod + emis : prod;
        } else {
            ok = replay_shift(C, res.s, F, f, mask);
        }
        if (!ok) {
            status_add(G, 2);
        } else {
            status_add(G, 0);
            if (fb(f.x) != fb(res.s.f.x) || fb(f.y) != fb(res.s.f.y) || fb(f.z) != fb(res.s.f.z)) status_add(G, 1);
            out = res.wSum * f;
            contract(F, res.wSum * g * mask, G, C.lds);
        }
    }
    if (G.replay) {
        G.replay[3 * (size_t)p + 0] = out.x;
        G.replay[3 * (size_t)p + 1] = out.y;
        G.replay[3 * (size_t)p + 2] = out.z;
    }
}

// One pixel of the path tracer (TraceCompletePath, path_gen.h:389-472) with detached pdfs.
__device__ __forceinline__ void adjoint_path_tracer(const AdjCtx& C, uint32_t p, uint32_t x, uint32_t y, V3 g) {
    const FrameArgs& A = C.A;
    const AdjArgs& G = C.G;
    const DScene& S = A.S;
    const cx_launch_params& lp = A.lp;
    Rng rng = rng_seed(tea64(p, frame_random(A)));
    const float jx = rng.f();
    const float jy = rng.f();
    float ux = ((float)x + jx) / (float)A.W;
    float uy = ((float)y + jy) / (float)A.H;
    ux = 2.0f * ux - 1.0f;
    uy = 2.0f * uy - 1.0f;
    const FrameCam cam = frame_cam(A);
    V3 pos = cam.pos;
    V3 dir = normalize(cam.U * ux + cam.V * uy + cam.W);
    Factors F;
    F.n = 0;
    F.overflow = false;
    V3 thr = v3(1.0f), radiance = v3(0.0f);
    const float inv1mNee = 1.0f / (1.0f - lp.neeProb);
    bool contributes = false;
    uint32_t depth = 0;
    for (uint32_t traceIdx = 0; traceIdx < (uint32_t)lp.maxPathLen; ++traceIdx) {
        Inter it;
        it.valid = false;
        trace_surface(S, C.tc, pos, dir, 1e-3f, 1e16f, it);
        if (!it.valid) break;
        ++depth;
        if (rng.f() < lp.neeProb) {
            thr = thr / lp.neeProb;
            NeeOut no;
            if (!nee_factor(C, it, rng, false, no)) break;   // occluded: radiance 0
            radiance = thr * no.brdf * no.color * no.distFactor;
            if (depth == 1) radiance = radiance + no.emis;
            Factor nf = no.fac;
            nf.v = nf.v / lp.neeProb;
            nf.dA = nf.dA / lp.neeProb;
            nf.dR = nf.dR / lp.neeProb;
            nf.dC = nf.dC / lp.neeProb;
            F.push(nf);
            contributes = true;
            break;
        }
        const BrdfSample bs = ggx_sample(S, it, rng);
        if (bs.pdf <= 0.0f) break;
        const BrdfD b = brdf_d(load_mat_from(G.evalMats, S, it), it.nrm, -it.dir, bs.outDir);
        pos = it.pos;
        dir = bs.outDir;
        thr = thr * (b.val / bs.pdf);
        thr = thr / (1.0f - lp.neeProb);
        F.push(Factor{b.val / bs.pdf * inv1mNee, b.dA / bs.pdf * inv1mNee, b.dR / bs.pdf * inv1mNee, v3(0.0f),
                      S.triMat[it.prim], NO_EMITTER});
    }
    V3 out = v3(0.0f);
    if (contributes && !F.overflow && !any_bad(radiance)) {
        status_add(G, 0);
        out = radiance;
        contract(F, g, G, C.lds);
    } else if (F.overflow) {
        status_add(G, 2);
    }
    if (G.replay) {
        G.replay[3 * (size_t)p + 0] = out.x;
        G.replay[3 * (size_t)p + 1] = out.y;
        G.replay[3 * (size_t)p + 2] = out.z;
    }
}

// With G.ldsSlots, a block sums its pixels' gradient terms in LDS (dynamic shared memory) and
// then adds each nonzero sum to the global arrays once: the terms of a frame fall on a few
// hundred addresses, and one global double add per term (about 10 M at 640 x 480) serialised
// on them (512² frame: 1.41 ms per adjoint, 0.29 ms of it the replay).  The sums keep the
// FP64 accumulation of the global adds; only their order differs, as it does between runs.
__global__ __launch_bounds__(ABLOCK) void k_adjoint(FrameArgs A, AdjArgs G) {
    __shared__ int s_stack[STACK_DEPTH * ABLOCK];
    extern __shared__ double s_grad[];
    const uint32_t slots = G.ldsSlots;
    for (uint32_t i = threadIdx.x; i < slots; i += ABLOCK) s_grad[i] = 0.0;
    if (slots) __syncthreads();
    TraceCtx tc;
    tc.stack = s_stack + threadIdx.x;
    tc.stride = ABLOCK;
    tc.closest = 0;
    tc.shadow = 0;
    const uint32_t p = blockIdx.x * ABLOCK + threadIdx.x;
    if (p < A.P) {
        const V3 g =
            G.dLdI ? V3{G.dLdI[3 * (size_t)p], G.dLdI[3 * (size_t)p + 1], G.dLdI[3 * (size_t)p + 2]} : v3(0.0f);
        const AdjCtx C{A, G, tc, slots ? s_grad : nullptr};
        if (A.lp.rendererType == CX_RESTIR_PT) adjoint_restir_pt(C, p, g);
        else adjoint_path_tracer(C, p, p % A.W, p / A.W, g);
    }
    if (!slots) return;
    __syncthreads();
    const uint32_t M = G.numMaterials;
    for (uint32_t i = threadIdx.x; i < slots; i += ABLOCK) {
        const double v = s_grad[i];
        if (v == 0.0) continue;
        double* dst = i < 3 * M ? G.gAlbedo + i : (i < 4 * M ? G.gRough + (i - 3 * M) : G.gEmit + (i - 4 * M));
        atomicAdd(dst, v);
    }
}

// Most LDS a block's gradient sums may take (4 M + 3 E doubles; 257 materials and 16 emitters
// need 8.6 KiB); larger tables add every term globally.
static constexpr size_t ADJ_LDS_MAX = 24 * 1024;

hipError_t launch_adjoint(const FrameArgs& A, const AdjArgs& G, hipStream_t s) {
    AdjArgs g = G;
    const size_t slots = 4 * (size_t)G.numMaterials + 3 * (size_t)G.numEmitters;
    g.ldsSlots = (G.gAlbedo || G.gRough || G.gEmit) && slots * sizeof(double) <= ADJ_LDS_MAX ? (uint32_t)slots : 0u;
    hipLaunchKernelGGL(k_adjoint, dim3((A.P + ABLOCK - 1) / ABLOCK), dim3(ABLOCK), g.ldsSlots * sizeof(double), s, A,
                       g);
    return hipGetLastError();
}

}  // namespace cx
