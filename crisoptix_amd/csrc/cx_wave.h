// Wavefront execution of the frame loop: path generation and hybrid shifts as explicit state
// machines whose contexts live in chunked SoA queues in HBM, and ray queries answered by
// separate traversal kernels that run at high occupancy.
//
// A machine step consumes the answer to the job's previous ray, runs the reference's logic up
// to the next trace call, and either appends a new ray to the next queue or finishes.  Closest-
// hit rays are appended from the front of a queue stripe and shadow rays from its back.
// Appends are wave-aggregated (one atomic per wavefront, ballot + prefix count), so a
// wavefront's surviving jobs land in consecutive slots and both the context stores and the
// next loads are coalesced.  A queue has 64 stripes with a counter pair each (wavefront w
// appends to stripe w mod 64): device-scope atomics on one address serialise at about 11 ns
// each on MI355X (372 us for the 32 k wavefronts of a 1080p pass), 64 addresses do not.
//
// Every machine computes exactly what the corresponding function of cx_restir_dev.h computes;
// only the points where control leaves for a traversal differ.  Traversal is a pure function
// of the ray, so the re-trace of path_gen.h:165-187 (same ray, same hit) is not issued again.
#pragma once

#include "cx_restir_dev.h"

namespace cx {

// ---- queues ----

static constexpr uint32_t EMIT_NONE = 0, EMIT_CLOSEST = 1, EMIT_SHADOW = 2;

static constexpr uint32_t QSTRIPES = 64;       // one per lane of the wavefront that scans them
static constexpr uint32_t QCOUNT_STRIDE = 16;  // words between stripes' counters (64 bytes)

struct Queue {
    uint4* ctx;         // [chunks][cap]
    uint32_t* job;      // [cap]
    float4* ray;        // [2][cap] origin.xyz tmin | dir.xyz tmax
    uint4* res;         // [cap] closest: prim t u v; shadow: occluded
    uint32_t* count;    // [QSTRIPES][QCOUNT_STRIDE]: closest count, shadow count of each stripe (one
                        // 64-bit word, closest in the low half)
    uint32_t* overflow; // bit 0: an append found its stripe full (sizing makes that impossible);
                        // bit 1: a tail kernel ended a job still running at its step bound
    uint32_t cap;       // QSTRIPES * stripeCap slots
    uint32_t stripeCap; // slots of a stripe: closest rays from its start, shadow rays from its end
    uint32_t jobs;      // most jobs the machine can have (sets the grid of its step kernels)
    uint32_t maxBlocks; // grid-stride kernels over this queue launch at most this many blocks
};

static constexpr uint32_t QUEUE_BLOCK = 128;   // threads per block of the queue kernels

// Blocks of a grid-stride launch over a machine with at most `jobs` jobs: enough for all of
// them, at most maxBlocks (the renderer's queue-grid setting).
CX_HD uint32_t queue_grid_blocks(uint32_t jobs, uint32_t maxBlocks) {
    const uint32_t need = (uint32_t)(((uint64_t)jobs + QUEUE_BLOCK - 1) / QUEUE_BLOCK);
    const uint32_t cap = maxBlocks > 0 ? maxBlocks : 1u;
    return need < 1u ? 1u : (need < cap ? need : cap);
}

// Slots per stripe for a machine of at most `jobs` jobs over P pixels: the most appends one
// stripe can receive.  Wavefront w appends to stripe w mod 64, at most 64 rays per append
// call.  A setup kernel (one thread per pixel, 2 wavefronts per block) makes ceil(jobs / P)
// append calls per lane; one with a thread per job (the final gather's shift starts) makes one
// call per lane over ceil(jobs / P) times the wavefronts, which the same bound covers; a step
// kernel makes one per loop pass of its grid-stride loop.
CX_HD uint32_t queue_stripe_cap(uint32_t jobs, uint32_t P, uint32_t maxBlocks) {
    auto cdiv = [](uint64_t a, uint64_t b) { return (a + b - 1) / b; };
    const uint64_t setupWaves = 2 * cdiv(P, QUEUE_BLOCK);
    const uint64_t perLane = cdiv(jobs, P > 0 ? P : 1);
    const uint64_t setup = cdiv(setupWaves, QSTRIPES) * 64 * perLane;
    const uint64_t nb = queue_grid_blocks(jobs, maxBlocks);
    const uint64_t step = cdiv(2 * nb, QSTRIPES) * 64 * cdiv(jobs, QUEUE_BLOCK * nb);
    uint64_t c = cdiv(jobs, QSTRIPES);
    if (setup > c) c = setup;
    if (step > c) c = step;
    return (uint32_t)c;
}

// Exclusive prefix sums of a queue's stripe counts, per block in LDS: the closest range of
// the input is stripe 0's closest rays, then stripe 1's, ...; the shadow range likewise.
struct QIndex {
    uint32_t pc[QSTRIPES + 1];
    uint32_t ps[QSTRIPES + 1];
};

// Every thread of the block calls it (one __syncthreads).
__device__ __forceinline__ void qindex_load(const Queue& q, QIndex& x) {
    if (threadIdx.x < 64) {
        const uint32_t lane = threadIdx.x;
        uint32_t c = q.count[lane * QCOUNT_STRIDE], s = q.count[lane * QCOUNT_STRIDE + 1];
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t yc = __shfl_up(c, off, 64), ys = __shfl_up(s, off, 64);
            if (lane >= (uint32_t)off) {
                c += yc;
                s += ys;
            }
        }
        x.pc[lane + 1] = c;
        x.ps[lane + 1] = s;
        if (lane == 0) {
            x.pc[0] = 0u;
            x.ps[0] = 0u;
        }
    }
    __syncthreads();
}

// The stripe s with pre[s] <= i < pre[s + 1] (pre non-decreasing, pre[0] = 0 <= i).
__device__ __forceinline__ uint32_t stripe_of(const uint32_t* pre, uint32_t i) {
    uint32_t s = 0;
#pragma unroll
    for (uint32_t step = QSTRIPES / 2; step > 0; step >>= 1)
        if (pre[s + step] <= i) s += step;
    return s;
}

struct Emit {
    uint32_t kind;
    V3 o, d;
    float tmin, tmax;
};

CX_HD Emit emit_none() { return Emit{EMIT_NONE, v3(0.0f), v3(0.0f), 0.0f, 0.0f}; }
CX_HD Emit emit_closest(V3 o, V3 d) { return Emit{EMIT_CLOSEST, o, d, 1e-3f, 1e16f}; }
CX_HD Emit emit_shadow(V3 o, V3 d, float tmax) { return Emit{EMIT_SHADOW, o, d, 1e-3f, tmax}; }

// Slot of the input queue handled by thread i: the closest range [0, nc) first, then the
// shadow range, each walked stripe by stripe.
__device__ __forceinline__ uint32_t in_slot(const Queue& q, const QIndex& x, uint32_t i) {
    const uint32_t nc = x.pc[QSTRIPES];
    if (i < nc) {
        const uint32_t s = stripe_of(x.pc, i);
        return s * q.stripeCap + (i - x.pc[s]);
    }
    const uint32_t j = i - nc;
    const uint32_t s = stripe_of(x.ps, j);
    return s * q.stripeCap + q.stripeCap - 1u - (j - x.ps[s]);
}

// Wave-aggregated append to the wavefront's stripe; every lane of the wavefront must call it.
// Returns the slot.  The stripe's closest and shadow counts are the two halves of one 64-bit
// word (closest low), advanced by one atomic, so every append sees both totals: a stripe whose
// two ranges would meet (never reached: the host sizes stripes for the most appends any stripe
// can receive) raises bit 0 of the overflow flag and wraps inside the stripe, so no access
// leaves the queue.
__device__ __forceinline__ uint32_t append(const Queue& q, uint32_t kind) {
    const uint64_t mc = __ballot(kind == EMIT_CLOSEST);
    const uint64_t ms = __ballot(kind == EMIT_SHADOW);
    const uint32_t lane = __lane_id();
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    const uint32_t stripe = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) % QSTRIPES;
    // 8-byte aligned: the counters start a 256-byte aligned allocation, QCOUNT_STRIDE words apart
    unsigned long long* cnt = (unsigned long long*)(q.count + stripe * QCOUNT_STRIDE);
    uint32_t bc = 0, bs = 0;
    if (lane == 0 && (mc | ms)) {
        const uint32_t nc = (uint32_t)__popcll(mc), ns = (uint32_t)__popcll(ms);
        const unsigned long long old = atomicAdd(cnt, (unsigned long long)nc | ((unsigned long long)ns << 32));
        bc = (uint32_t)old;
        bs = (uint32_t)(old >> 32);
        if ((uint64_t)bc + nc + (uint64_t)bs + ns > q.stripeCap) atomicOr(q.overflow, 1u);
    }
    bc = __shfl(bc, 0, 64);
    bs = __shfl(bs, 0, 64);
    const uint32_t base = stripe * q.stripeCap;
    if (kind == EMIT_CLOSEST) return base + (bc + (uint32_t)__popcll(mc & lt)) % q.stripeCap;
    if (kind == EMIT_SHADOW) return base + q.stripeCap - 1u - (bs + (uint32_t)__popcll(ms & lt)) % q.stripeCap;
    return 0xFFFFFFFFu;
}

__device__ __forceinline__ void put_ray(const Queue& q, uint32_t slot, const Emit& e) {
    q.ray[slot] = float4{e.o.x, e.o.y, e.o.z, e.tmin};
    q.ray[(size_t)q.cap + slot] = float4{e.d.x, e.d.y, e.d.z, e.tmax};
}

__device__ __forceinline__ void get_ray(const Queue& q, uint32_t slot, V3& o, V3& d) {
    const float4 a = q.ray[slot], b = q.ray[(size_t)q.cap + slot];
    o = V3{a.x, a.y, a.z};
    d = V3{b.x, b.y, b.z};
}

template <int N>
struct Ctx {
    uint4 w[N];
};

template <int N>
__device__ __forceinline__ Ctx<N> load_ctx(const Queue& q, uint32_t slot) {
    Ctx<N> c;
#pragma unroll
    for (int i = 0; i < N; i++) c.w[i] = q.ctx[(size_t)i * q.cap + slot];
    return c;
}

template <int N>
__device__ __forceinline__ void store_ctx(const Queue& q, uint32_t slot, const Ctx<N>& c) {
#pragma unroll
    for (int i = 0; i < N; i++) q.ctx[(size_t)i * q.cap + slot] = c.w[i];
}

// Interaction from a closest-hit answer (__closesthit__mesh / __miss__main): the normal is
// normalised; a miss only clears `valid` and sets the direction.
__device__ __forceinline__ Inter inter_from_hit(const DScene& S, uint4 h, V3 d) {
    Inter it;
    it.dir = d;
    it.prim = h.x;
    it.bu = bf(h.z);
    it.bv = bf(h.w);
    it.valid = h.x != INVALID_PRIM;
    it.pos = v3(0.0f);
    it.nrm = v3(0.0f);
    if (!it.valid) return it;
    const uint32_t i0 = S.idx[3 * h.x], i1 = S.idx[3 * h.x + 1], i2 = S.idx[3 * h.x + 2];
    it.pos = interp(it.bu, it.bv, load_v3(S.pos, i0), load_v3(S.pos, i1), load_v3(S.pos, i2));
    it.nrm = normalize(interp(it.bu, it.bv, load_v3(S.nrm, i0), load_v3(S.nrm, i1), load_v3(S.nrm, i2)));
    return it;
}

// Reconstructs a stored interaction: rehydration, plus normalisation when it came from a trace.
__device__ __forceinline__ Inter inter_restore(const DScene& S, const Packed& p, bool traced) {
    Inter it = rehydrate(S, p);
    if (traced && it.valid) it.nrm = normalize(it.nrm);
    return it;
}

// ---- packing helpers: contexts are written and read as a sequence of 32-bit words ----

template <int N>
struct Wr {
    uint32_t a[4 * N];
    int n;
    __device__ __forceinline__ Wr() : n(0) {
#pragma unroll
        for (int i = 0; i < 4 * N; i++) a[i] = 0u;
    }
    __device__ __forceinline__ void u(uint32_t x) { a[n++] = x; }
    __device__ __forceinline__ void f(float x) { a[n++] = fb(x); }
    __device__ __forceinline__ void v(V3 x) { f(x.x); f(x.y); f(x.z); }
    __device__ __forceinline__ void u64(uint64_t x) { u((uint32_t)x); u((uint32_t)(x >> 32)); }
    __device__ __forceinline__ void pk(const Packed& p) { u(p.prim); f(p.bu); f(p.bv); v(p.dir); }
    __device__ __forceinline__ Ctx<N> ctx() const {
        Ctx<N> c;
#pragma unroll
        for (int i = 0; i < N; i++) c.w[i] = uint4{a[4 * i], a[4 * i + 1], a[4 * i + 2], a[4 * i + 3]};
        return c;
    }
};

template <int N>
struct Rd {
    uint32_t a[4 * N];
    int n;
    __device__ __forceinline__ explicit Rd(const Ctx<N>& c) : n(0) {
#pragma unroll
        for (int i = 0; i < N; i++) {
            a[4 * i] = c.w[i].x;
            a[4 * i + 1] = c.w[i].y;
            a[4 * i + 2] = c.w[i].z;
            a[4 * i + 3] = c.w[i].w;
        }
    }
    __device__ __forceinline__ uint32_t u() { return a[n++]; }
    __device__ __forceinline__ float f() { return bf(a[n++]); }
    __device__ __forceinline__ V3 v() {
        const float x = f(), y = f(), z = f();
        return V3{x, y, z};
    }
    __device__ __forceinline__ uint64_t u64() {
        const uint64_t lo = u(), hi = u();
        return lo | (hi << 32);
    }
    __device__ __forceinline__ Packed pk() {
        Packed p;
        p.prim = u();
        p.bu = f();
        p.bv = f();
        p.dir = v();
        return p;
    }
};

// ======================================================================================
// PrefixGen: TracePrefix (path_gen.h:74-250)
// ======================================================================================

static constexpr uint32_t PG_TRACE = 1, PG_NEE = 2;
// The queued context of a walk (11 chunks, 176 bytes): the prefix record, the generator, the
// state word, and the traced interaction and emitter colour of a pending NEE ray.  The rest of
// PgState is rebuilt from the job's queued ray (pg_unpack): while a closest-hit ray is pending,
// cur and dir are its origin and direction; while an NEE ray is pending, lightDir and dist are its
// direction and tmax, and dir (the direction that reached the interaction) is the interaction's
// stored incoming direction.
static constexpr int PG_CHUNKS = 11;

struct PgState {
    Prefix pre;
    uint64_t rng;
    V3 cur, dir;
    uint32_t traceIdx, postRecon, state;
    Packed it;            // traced interaction (normalised normal)
    V3 lightDir;
    float dist;
    V3 color;
};

__device__ __forceinline__ Ctx<PG_CHUNKS> pg_pack(const PgState& s) {
    Wr<PG_CHUNKS> w;
    const Prefix& p = s.pre;
    w.u64(p.rng); w.u(p.flags); w.f(p.p); w.v(p.f); w.f(p.pathLen); w.v(p.postReconF); w.v(p.reconOutDir);
    w.pk(p.primary); w.pk(p.recon); w.pk(p.last);
    w.u64(s.rng); w.u(s.traceIdx | (s.postRecon << 8) | (s.state << 16));
    w.pk(s.it); w.v(s.color);
    return w.ctx();
}

// ray: the job's queued ray {origin, tmin}, {direction, tmax}.
__device__ __forceinline__ PgState pg_unpack(const Ctx<PG_CHUNKS>& c, float4 ra, float4 rb) {
    Rd<PG_CHUNKS> r(c);
    PgState s;
    Prefix& p = s.pre;
    p.rng = r.u64(); p.flags = r.u(); p.p = r.f(); p.f = r.v(); p.pathLen = r.f(); p.postReconF = r.v();
    p.reconOutDir = r.v(); p.primary = r.pk(); p.recon = r.pk(); p.last = r.pk();
    s.rng = r.u64();
    const uint32_t m = r.u();
    s.traceIdx = m & 0xFFu;
    s.postRecon = (m >> 8) & 1u;
    s.state = m >> 16;
    s.it = r.pk(); s.color = r.v();
    s.cur = V3{ra.x, ra.y, ra.z};
    if (s.state == PG_NEE) {
        s.dir = s.it.dir;
        s.lightDir = V3{rb.x, rb.y, rb.z};
        s.dist = rb.w;
    } else {
        s.dir = V3{rb.x, rb.y, rb.z};
        s.lightDir = v3(0.0f);
        s.dist = 0.0f;
    }
    return s;
}

__device__ __forceinline__ uint32_t pg_max_len(const cx_launch_params& lp) {
    return is_pt(lp) ? (uint32_t)lp.maxPathLen : (uint32_t)lp.prefixLen;
}

// The generator fields the init kernels set before pg_start.
__device__ __forceinline__ void pg_reset(PgState& s, uint64_t rng) {
    s.rng = rng;
    s.pre.primary = packed_invalid();
    s.pre.recon = packed_invalid();
    s.pre.last = packed_invalid();
    s.pre.reconOutDir = v3(0.0f);
    s.lightDir = v3(0.0f);
    s.dist = 0.0f;
    s.color = v3(0.0f);
}

// Start of TracePrefix with the camera ray: returns the first emission.
__device__ __forceinline__ Emit pg_start(const cx_launch_params& lp, PgState& s, V3 origin, V3 dir) {
    s.pre.flags = 0u;
    s.pre.pathLen = 0.0f;
    s.pre.rng = s.rng;
    s.pre.p = 1.0f;
    s.pre.f = v3(1.0f);
    s.pre.postReconF = v3(1.0f);
    s.pre.set_valid(true);
    s.cur = origin;
    s.dir = dir;
    s.traceIdx = 0;
    s.postRecon = 0;
    s.it = packed_invalid();
    if (pg_max_len(lp) == 0) {   // loop body never runs: lastInt stays the default (invalid)
        s.pre.last = packed_invalid();
        return emit_none();
    }
    s.state = PG_TRACE;
    return emit_closest(origin, dir);
}

// After the closest hit of the current vertex (path_gen.h:100-247).
__device__ __forceinline__ Emit pg_on_hit(const DScene& S, const cx_launch_params& lp, PgState& s, const Inter& it) {
    Prefix& pre = s.pre;
    Rng rng;
    rng.s = s.rng;
    const uint32_t maxLen = pg_max_len(lp);
    if (!it.valid) {
        pre.set_valid(false);
        return emit_none();
    }
    pre.pathLen += distance(s.cur, it.pos);
    pre.set_length(pre.length() + 1);
    if (pre.length() == 1) pre.primary = it.pack();
    if (is_cond(lp) && s.traceIdx == maxLen - 1) {
        if (!s.postRecon) {
            pre.recon = it.pack();
            pre.set_recon_idx(pre.length());
            pre.reconOutDir = v3(0.0f);
        }
        pre.set_valid(true);
        pre.last = it.pack();
        return emit_none();
    }
    const bool forceNee = is_pt(lp) && s.traceIdx == maxLen - 1;
    if (rng.f() < lp.neeProb || forceNee) {
        pre.p *= lp.neeProb;
        const EmitterSample es = sample_emitter(S, rng);
        s.lightDir = normalize(es.pos - it.pos);
        s.dist = length(es.pos - it.pos);
        s.color = es.color;
        s.it = it.pack();
        s.rng = rng.s;
        s.state = PG_NEE;
        return emit_shadow(it.pos, s.lightDir, s.dist);
    }
    const BrdfSample bs = ggx_sample(S, it, rng);
    s.rng = rng.s;
    if (bs.pdf <= 0.0f) {
        pre.set_valid(false);
        return emit_none();
    }
    const float reconDistance = distance(s.cur, it.pos);
    const bool canRecon =
        reconDistance > lp.reconMinDistance && bs.roughness > lp.reconMinRoughness && pre.length() >= 2;
    if (!s.postRecon && canRecon) {
        pre.recon = it.pack();
        pre.set_recon_idx(pre.length());
        pre.reconOutDir = bs.outDir;
        s.postRecon = 1;
    }
    s.cur = it.pos;
    s.dir = bs.outDir;
    pre.f = pre.f * bs.brdf;
    pre.p *= bs.pdf * (1.0f - lp.neeProb);
    if (s.postRecon) pre.postReconF = pre.postReconF * bs.brdf;
    s.traceIdx++;
    if (s.traceIdx < maxLen) {
        s.state = PG_TRACE;
        return emit_closest(s.cur, s.dir);
    }
    pre.last = it.pack();
    return emit_none();
}

// After the NEE shadow ray (path_gen.h:158-210).
__device__ __forceinline__ Emit pg_on_shadow(const DScene& S, const cx_launch_params& lp, PgState& s, bool occluded) {
    Prefix& pre = s.pre;
    if (occluded) {
        pre.set_valid(false);
        return emit_none();
    }
    const Inter it = inter_restore(S, s.it, true);
    if (!s.postRecon) {   // the re-trace of the same ray returns this same hit
        pre.recon = s.it;
        pre.set_recon_idx(pre.length());
        pre.reconOutDir = s.dir;
        s.postRecon = 1;
    }
    const BrdfEval be = ggx_eval(S, it, s.lightDir);
    if (be.pdf <= 0.0f) {
        pre.set_valid(false);
        return emit_none();
    }
    const float distFactor = 1.0f / (s.dist * s.dist);
    pre.f = pre.f * (be.brdf * s.color * distFactor);
    if (s.postRecon) pre.postReconF = pre.postReconF * (be.brdf * s.color * distFactor);
    if (pre.length() == 1) pre.f = pre.f + be.emission * distFactor;
    pre.set_nee(true);
    pre.set_valid(true);
    pre.last = s.it;
    return emit_none();
}

// ======================================================================================
// SuffixGen: TraceSuffix (path_gen.h:252-387)
// ======================================================================================

static constexpr uint32_t SG_TRACE = 1, SG_NEE = 2, SG_RECON = 3;
// The queued context of a suffix walk (10 chunks): the suffix record, the generator, the state
// words and the emitter colour, then 6 words whose use depends on the pending ray: the NEE ray's
// interaction (SG_NEE), or the BSDF value and the light distance (SG_RECON).  The rest of SgState
// is rebuilt from the job's queued ray (sg_unpack): lastDir is a closest-hit ray's direction
// (SG_TRACE), lightDir the NEE or reconnection ray's direction, lightLen the NEE ray's tmax.
static constexpr int SG_CHUNKS = 10;

struct SgState {
    Suffix suf;
    uint64_t rng;
    uint32_t vertIdx, maxLen, prefixLen, postRecon, state, itTraced;
    Packed it;
    V3 lastDir;
    V3 lightDir;
    float lightLen;
    V3 color;
    V3 brdf;
};

__device__ __forceinline__ Ctx<SG_CHUNKS> sg_pack(const SgState& s) {
    Wr<SG_CHUNKS> w;
    const Suffix& u = s.suf;
    w.u64(u.rng); w.u(u.flags); w.f(u.p); w.v(u.f); w.v(u.postReconF); w.v(u.reconOutDir);
    w.pk(u.recon); w.pk(u.lastPrefix);
    w.u64(s.rng);
    w.u(s.vertIdx | (s.maxLen << 8) | (s.prefixLen << 16) | (s.postRecon << 24) | (s.itTraced << 25));
    w.u(s.state);
    w.v(s.color);
    if (s.state == SG_RECON) {
        w.v(s.brdf); w.f(s.lightLen);
    } else {
        w.pk(s.it);
    }
    return w.ctx();
}

// ra, rb: the job's queued ray {origin, tmin}, {direction, tmax}.
__device__ __forceinline__ SgState sg_unpack(const Ctx<SG_CHUNKS>& c, float4 ra, float4 rb) {
    Rd<SG_CHUNKS> r(c);
    SgState s;
    Suffix& u = s.suf;
    u.rng = r.u64(); u.flags = r.u(); u.p = r.f(); u.f = r.v(); u.postReconF = r.v(); u.reconOutDir = r.v();
    u.recon = r.pk(); u.lastPrefix = r.pk();
    s.rng = r.u64();
    const uint32_t m = r.u();
    s.vertIdx = m & 0xFFu;
    s.maxLen = (m >> 8) & 0xFFu;
    s.prefixLen = (m >> 16) & 0xFFu;
    s.postRecon = (m >> 24) & 1u;
    s.itTraced = (m >> 25) & 1u;
    s.state = r.u();
    s.color = r.v();
    const V3 d = V3{rb.x, rb.y, rb.z};
    s.lastDir = d;
    s.lightDir = d;
    s.brdf = v3(0.0f);
    s.lightLen = rb.w;
    if (s.state == SG_RECON) {
        s.brdf = r.v(); s.lightLen = r.f();
        s.it = packed_invalid();
    } else {
        s.it = r.pk();
    }
    (void)ra;
    return s;
}

// One vertex of the suffix loop once its interaction is known (path_gen.h:296-385).
__device__ __forceinline__ Emit sg_vertex(const DScene& S, const cx_launch_params& lp, SgState& s, const Inter& it,
                                          V3 lastPos) {
    Suffix& suf = s.suf;
    Rng rng;
    rng.s = s.rng;
    const bool forceNee = s.vertIdx == s.maxLen - 1;
    if (rng.f() < lp.neeProb || forceNee) {
        if (!forceNee) suf.p *= lp.neeProb;
        const EmitterSample em = sample_emitter(S, rng);
        s.rng = rng.s;
        const V3 lightVec = em.pos - it.pos;
        s.lightLen = length(lightVec);
        s.lightDir = normalize(lightVec);
        if (any_bad(s.lightDir)) return emit_none();
        s.color = em.color;
        s.it = it.pack();
        s.state = SG_NEE;
        return emit_shadow(it.pos, s.lightDir, s.lightLen);
    }
    const BrdfSample bs = ggx_sample(S, it, rng);
    s.rng = rng.s;
    if (bs.pdf <= 0.0f) return emit_none();
    suf.p *= bs.pdf * (1.0f - lp.neeProb);
    suf.f = suf.f * bs.brdf;
    if (s.postRecon) suf.postReconF = suf.postReconF * bs.brdf;
    const float reconDistance = distance(lastPos, it.pos);
    const bool canRecon =
        reconDistance > lp.reconMinDistance && bs.roughness > lp.reconMinRoughness && s.prefixLen >= 1;
    if (!s.postRecon && canRecon) {
        suf.recon = it.pack();
        suf.set_recon_idx(s.prefixLen);   // path_gen.h:378
        suf.reconOutDir = bs.outDir;
        s.postRecon = 1;
    }
    s.lastDir = bs.outDir;
    suf.set_length(suf.length() + 1);
    s.vertIdx++;
    if (s.vertIdx < s.maxLen) {
        s.state = SG_TRACE;
        return emit_closest(it.pos, s.lastDir);
    }
    return emit_none();
}

// Start of TraceSuffix from a prefix (path_gen.h:258-294).
__device__ __forceinline__ Emit sg_start(const DScene& S, const cx_launch_params& lp, SgState& s, const Prefix& pre) {
    Suffix& suf = s.suf;
    suf.f = v3(1.0f);
    suf.p = 1.0f;
    suf.postReconF = v3(1.0f);
    suf.rng = s.rng;
    suf.lastPrefix = pre.last;
    suf.flags = 0u;
    s.postRecon = 0;
    s.vertIdx = 0;
    s.lastDir = v3(0.0f);
    s.state = 0;
    s.itTraced = 0;
    s.prefixLen = pre.length();
    if (!pre.valid() || pre.nee()) return emit_none();
    const uint32_t maxLen = (uint32_t)lp.maxPathLen - pre.length();
    s.maxLen = maxLen & 0xFFu;
    const Inter it = rehydrate(S, pre.last);
    if (!it.valid) return emit_none();
    if (maxLen == 0) return emit_none();
    return sg_vertex(S, lp, s, it, it.pos);
}

__device__ __forceinline__ Emit sg_on_hit(const DScene& S, const cx_launch_params& lp, SgState& s, const Inter& it,
                                          V3 rayOrigin) {
    s.itTraced = 1;
    if (!it.valid) return emit_none();
    return sg_vertex(S, lp, s, it, rayOrigin);
}

__device__ __forceinline__ Emit sg_finish(SgState& s) {
    Suffix& suf = s.suf;
    const float distFactor = 1.0f / (s.lightLen * s.lightLen);
    suf.f = suf.f * (s.brdf * s.color * distFactor);
    suf.postReconF = suf.postReconF * (s.color * distFactor);
    if (suf.recon_idx() < suf.length()) suf.postReconF = suf.postReconF * s.brdf;
    suf.set_valid(true);
    return emit_none();
}

__device__ __forceinline__ Emit sg_on_shadow(const DScene& S, SgState& s, bool occluded) {
    if (occluded) return emit_none();
    const Inter it = inter_restore(S, s.it, s.itTraced != 0);
    const BrdfEval be = ggx_eval(S, it, s.lightDir);
    if (be.pdf <= 0.0f) return emit_none();
    s.suf.set_length(s.suf.length() + 1);
    s.brdf = be.brdf;
    if (!s.postRecon) {
        s.state = SG_RECON;
        return emit_closest(it.pos, s.lightDir);
    }
    return sg_finish(s);
}

__device__ __forceinline__ Emit sg_on_recon(SgState& s, const Inter& ri) {
    if (!ri.valid) return emit_none();
    s.suf.recon = ri.pack();
    s.suf.set_recon_idx(s.suf.length());
    s.suf.reconOutDir = v3(0.0f);
    s.postRecon = 1;
    return sg_finish(s);
}

// ======================================================================================
// CompletePath: TraceCompletePath (path_gen.h:389-472)
// ======================================================================================

static constexpr uint32_t CP_TRACE = 1, CP_NEE = 2;
static constexpr int CP_CHUNKS = 7;

struct CpState {
    uint64_t rng;
    uint32_t traceIdx, depth, state;
    V3 cur, dir, thr, radiance;
    Packed it;
    V3 lightDir;
    float dist;
    V3 color;
};

__device__ __forceinline__ Ctx<CP_CHUNKS> cp_pack(const CpState& s) {
    Wr<CP_CHUNKS> w;
    w.u64(s.rng); w.u(s.traceIdx | (s.depth << 8) | (s.state << 16));
    w.v(s.cur); w.v(s.dir); w.v(s.thr); w.pk(s.it); w.v(s.lightDir); w.f(s.dist); w.v(s.color);
    return w.ctx();
}

__device__ __forceinline__ CpState cp_unpack(const Ctx<CP_CHUNKS>& c) {
    Rd<CP_CHUNKS> r(c);
    CpState s;
    s.rng = r.u64();
    const uint32_t m = r.u();
    s.traceIdx = m & 0xFFu;
    s.depth = (m >> 8) & 0xFFu;
    s.state = m >> 16;
    s.cur = r.v(); s.dir = r.v(); s.thr = r.v(); s.it = r.pk(); s.lightDir = r.v(); s.dist = r.f(); s.color = r.v();
    s.radiance = v3(0.0f);
    return s;
}

// The state of a complete path before its first trace.
__device__ __forceinline__ void cp_reset(CpState& s, uint64_t rng, V3 origin, V3 dir) {
    s.rng = rng;
    s.traceIdx = 0;
    s.depth = 0;
    s.state = CP_TRACE;
    s.cur = origin;
    s.dir = dir;
    s.thr = v3(1.0f);
    s.radiance = v3(0.0f);
    s.it = packed_invalid();
    s.lightDir = v3(0.0f);
    s.dist = 0.0f;
    s.color = v3(0.0f);
}

// ---- fresh contexts: a walk before its first trace is a function of its generator state and
// its first ray, which the queue holds.  An init kernel stores only chunk 0 = {rng, CTX_FRESH}
// (word 2 holds the prefix flags or the machine's state word, where bit 31 is never set) and
// the first step rebuilds the rest, instead of a full context write and read per pixel. ----
static constexpr uint32_t CTX_FRESH = 0x80000000u;

__device__ __forceinline__ void store_fresh(const Queue& q, uint32_t slot, uint64_t rng) {
    q.ctx[slot] = uint4{(uint32_t)rng, (uint32_t)(rng >> 32), CTX_FRESH, 0u};
}

__device__ __forceinline__ PgState pg_load(const Queue& q, uint32_t slot, const cx_launch_params& lp) {
    const uint4 c0 = q.ctx[slot];
    PgState s;
    if (c0.z & CTX_FRESH) {
        V3 o, d;
        get_ray(q, slot, o, d);
        pg_reset(s, (uint64_t)c0.x | ((uint64_t)c0.y << 32));
        (void)pg_start(lp, s, o, d);   // re-emits the queued camera ray
        return s;
    }
    Ctx<PG_CHUNKS> c;
    c.w[0] = c0;
#pragma unroll
    for (int i = 1; i < PG_CHUNKS; i++) c.w[i] = q.ctx[(size_t)i * q.cap + slot];
    return pg_unpack(c, q.ray[slot], q.ray[(size_t)q.cap + slot]);
}

__device__ __forceinline__ CpState cp_load(const Queue& q, uint32_t slot) {
    const uint4 c0 = q.ctx[slot];
    CpState s;
    if (c0.z & CTX_FRESH) {
        V3 o, d;
        get_ray(q, slot, o, d);
        cp_reset(s, (uint64_t)c0.x | ((uint64_t)c0.y << 32), o, d);
        return s;
    }
    Ctx<CP_CHUNKS> c;
    c.w[0] = c0;
#pragma unroll
    for (int i = 1; i < CP_CHUNKS; i++) c.w[i] = q.ctx[(size_t)i * q.cap + slot];
    return cp_unpack(c);
}

// Returns an emission; when none, `done` holds the radiance in s.radiance.
__device__ __forceinline__ Emit cp_on_hit(const DScene& S, const cx_launch_params& lp, CpState& s, const Inter& it) {
    if (!it.valid) {
        s.radiance = v3(0.0f);
        return emit_none();
    }
    ++s.depth;
    Rng rng;
    rng.s = s.rng;
    if (rng.f() < lp.neeProb) {
        s.thr = s.thr / lp.neeProb;
        const EmitterSample es = sample_emitter(S, rng);
        s.rng = rng.s;
        s.lightDir = normalize(es.pos - it.pos);
        s.dist = length(es.pos - it.pos);
        s.color = es.color;
        s.it = it.pack();
        s.state = CP_NEE;
        return emit_shadow(it.pos, s.lightDir, s.dist);
    }
    const BrdfSample bs = ggx_sample(S, it, rng);
    s.rng = rng.s;
    if (bs.pdf <= 0.0f) {
        s.radiance = v3(0.0f);
        return emit_none();
    }
    s.cur = it.pos;
    s.dir = bs.outDir;
    s.thr = s.thr * (bs.brdf / bs.pdf);
    s.thr = s.thr / (1.0f - lp.neeProb);
    s.traceIdx++;
    if (s.traceIdx < (uint32_t)lp.maxPathLen) {
        s.state = CP_TRACE;
        return emit_closest(s.cur, s.dir);
    }
    s.radiance = v3(0.0f);
    return emit_none();
}

__device__ __forceinline__ Emit cp_on_shadow(const DScene& S, CpState& s, bool occluded) {
    if (occluded) {
        s.radiance = v3(0.0f);
        return emit_none();
    }
    const Inter it = inter_restore(S, s.it, true);
    const BrdfEval be = ggx_eval(S, it, s.lightDir);
    const float distFactor = 1.0f / (s.dist * s.dist);
    s.radiance = s.thr * be.brdf * s.color * distFactor;
    if (s.depth == 1) s.radiance = s.radiance + be.emission;
    return emit_none();
}

// ======================================================================================
// PrefixFG: TracePrefixForFinalGather (path_gen.h:474-531)
// ======================================================================================

static constexpr int FG_CHUNKS = 3;

struct FgState {
    uint64_t rng;
    uint32_t idx;
    V3 thr;
    float p;
    V3 dir;
};

__device__ __forceinline__ Ctx<FG_CHUNKS> fg_pack(const FgState& s) {
    Wr<FG_CHUNKS> w;
    w.u64(s.rng); w.u(s.idx); w.f(s.p); w.v(s.thr); w.v(s.dir);
    return w.ctx();
}

__device__ __forceinline__ FgState fg_unpack(const Ctx<FG_CHUNKS>& c) {
    Rd<FG_CHUNKS> r(c);
    FgState s;
    s.rng = r.u64(); s.idx = r.u(); s.p = r.f(); s.thr = r.v(); s.dir = r.v();
    return s;
}

// ======================================================================================
// Shift: hybrid shift of a prefix (prefix_reuse.h:10-99) or a suffix (suffix_reuse.h:10-123)
// ======================================================================================

// Where a shift reads its source / destination record.
static constexpr uint32_t REF_PREFIX_FRONT = 0, REF_PREFIX_BACK = 1, REF_PREFIX_SCRATCH = 2, REF_SUFFIX_FRONT = 3,
                          REF_SUFFIX_BACK = 4, REF_SUFFIX_SCRATCH = 5, REF_PACKED_TRACED = 6;
CX_HD uint32_t make_ref(uint32_t buf, uint32_t pixel) { return (buf << 28) | pixel; }

struct RecordSet {
    const uint4* prefixRes[2];
    const uint4* suffixRes[2];
    const uint4* prefixScratch;   // canonical prefix records (PrefixRes layout)
    const uint4* suffixScratch;   // canonical suffix records (SuffixRes layout)
    const uint4* packedScratch;   // [P] x 2 chunks: Packed of a traced interaction
    const uint4* suffixFrontAos;  // record-major copy of suffixRes[front], or null (final gather only)
    uint32_t front, back, P;
};

__device__ __forceinline__ Prefix load_prefix_ref(const RecordSet& R, uint32_t ref) {
    const uint32_t buf = ref >> 28, p = ref & 0x0FFFFFFFu;
    const uint4* b = buf == REF_PREFIX_FRONT ? R.prefixRes[R.front]
                                             : (buf == REF_PREFIX_BACK ? R.prefixRes[R.back] : R.prefixScratch);
    return load_prefix(b, R.P, p).s;
}

__device__ __forceinline__ Suffix load_suffix_ref(const RecordSet& R, uint32_t ref) {
    const uint32_t buf = ref >> 28, p = ref & 0x0FFFFFFFu;
    if (buf == REF_SUFFIX_FRONT && R.suffixFrontAos) return load_suffix_aos(R.suffixFrontAos, p).s;
    const uint4* b = buf == REF_SUFFIX_FRONT ? R.suffixRes[R.front]
                                             : (buf == REF_SUFFIX_BACK ? R.suffixRes[R.back] : R.suffixScratch);
    return load_suffix(b, R.P, p).s;
}

// Destination interaction of a suffix shift: the destination suffix's last prefix vertex
// (rehydrated, suffix_reuse.h:217) or a traced interaction (final_gather.cu:265).
__device__ __forceinline__ Inter load_suffix_dst(const DScene& S, const RecordSet& R, uint32_t ref) {
    const uint32_t buf = ref >> 28, p = ref & 0x0FFFFFFFu;
    if (buf == REF_PACKED_TRACED) {
        const uint4 a = R.packedScratch[2 * (size_t)p], b = R.packedScratch[2 * (size_t)p + 1];
        const Packed pk = Packed{a.x, bf(a.y), bf(a.z), V3{bf(a.w), bf(b.x), bf(b.y)}};
        return inter_restore(S, pk, true);
    }
    return rehydrate(S, load_suffix_ref(R, ref).lastPrefix);
}

// Everything a wavefront kernel reads or writes.  qin is the queue whose rays were just
// answered, qout the queue the kernel appends to; both share one allocation per side and are
// viewed with the capacity and chunk count of the machine in flight.
struct WaveArgs {
    FrameArgs A;
    RecordSet R;
    Queue qin, qout;
    uint4* prefixScratch;   // [PREFIX_CHUNKS][P] canonical prefix of the pass
    uint4* suffixScratch;   // [SUFFIX_CHUNKS][P] canonical suffix of the pass
    uint4* packedScratch;   // [P][2] last vertex of a final-gather prefix (Packed)
    uint4* suffixAos;       // [P][SUFFIX_AOS_STRIDE] record-major copy of the front suffix reservoirs
    uint64_t* rngScratch;   // [P] pixel generator after path generation / neighbour selection
    float4* shiftRes;       // [jobs] f.xyz, Jacobian of each finished shift
    uint32_t* pixFlags;     // [P] per-pass pixel flags
    uint32_t* validMask;    // [P] valid-neighbour mask of the final gather
    // The final gather's querying pixels in Morton order of their query points: keys [P] (per
    // pixel, 0xFFFFFFFF where it does not query), then sorted with the pixel as value into
    // gatherList[0, *gatherCount); the Alt arrays and sortTemp are the sort's scratch.
    uint32_t* gatherKeys;
    uint32_t* gatherList;
    uint32_t* gatherKeysAlt;
    uint32_t* gatherListAlt;
    uint32_t* gatherCount;
    void* gatherSortTemp;
    size_t gatherSortBytes;
    float4* gatherPos;      // [P] query point of each querying pixel
    V3 mortonLo, mortonScale;   // quantisation of the query points (the scene's bounds)
    uint32_t* nbr;          // [P][16] neighbour pixel of each spatial slot
    float4* radScratch;     // [P] radiance summed over the final-gather prefixes
    float4* fgScratch;      // [P] final-gather prefix throughput xyz, pdf
    float4* camDir;         // [P] camera ray direction of the final gather
    uint32_t S;             // spatial neighbour slots of the pass
    uint32_t gIdx;          // final-gather prefix index (0 = canonical)
};

static constexpr uint32_t SH_TRACE = 1, SH_SHADOW = 2;
static constexpr int SH_CHUNKS = 4;

struct ShState {
    uint32_t srcRef, isSuffix, idx, count, state, currTraced;
    uint64_t srng;
    V3 thr;
    Packed curr;
};

__device__ __forceinline__ Ctx<SH_CHUNKS> sh_pack(const ShState& s) {
    Wr<SH_CHUNKS> w;
    w.u(s.srcRef);
    w.u(s.idx | (s.count << 8) | (s.state << 16) | (s.isSuffix << 24) | (s.currTraced << 25));
    w.u64(s.srng); w.v(s.thr); w.pk(s.curr);
    return w.ctx();
}

__device__ __forceinline__ ShState sh_unpack(const Ctx<SH_CHUNKS>& c) {
    Rd<SH_CHUNKS> r(c);
    ShState s;
    s.srcRef = r.u();
    const uint32_t m = r.u();
    s.idx = m & 0xFFu;
    s.count = (m >> 8) & 0xFFu;
    s.state = (m >> 16) & 0xFFu;
    s.isSuffix = (m >> 24) & 1u;
    s.currTraced = (m >> 25) & 1u;
    s.srng = r.u64(); s.thr = r.v(); s.curr = r.pk();
    return s;
}

// Replay step or reconnection ray, from the current interaction.
__device__ __forceinline__ Emit sh_continue(const DScene& S, const RecordSet& R, ShState& s, const Inter& curr) {
    if (s.idx < s.count) {
        Rng srng;
        srng.s = s.srng;
        srng.f();   // dummy draw for the NEE decision (prefix_reuse.h:44, suffix_reuse.h:43)
        const BrdfSample bs = ggx_sample(S, curr, srng);
        s.srng = srng.s;
        if (bs.pdf <= 0.0f) return emit_none();
        s.thr = s.thr * bs.brdf;
        s.curr = curr.pack();
        s.state = SH_TRACE;
        return emit_closest(curr.pos, bs.outDir);
    }
    // an if, not ?: over the two records: the conditional operator on whole records made the
    // compiler load both into private memory and read the field back from there
    Packed reconP;
    if (s.isSuffix) reconP = load_suffix_ref(R, s.srcRef).recon;
    else reconP = load_prefix_ref(R, s.srcRef).recon;
    const Inter srcRecon = rehydrate(S, reconP);
    const V3 reconVec = srcRecon.pos - curr.pos;
    const float reconLen = length(reconVec);
    const V3 reconDir = normalize(reconVec);
    if (!s.isSuffix && isinf(reconLen)) return emit_none();
    if (any_bad(reconDir)) return emit_none();
    s.curr = curr.pack();
    s.state = SH_SHADOW;
    return emit_shadow(curr.pos, reconDir, reconLen);
}

// Start of a shift: returns the first emission (none = finished with f = 0, J = 0).
__device__ __forceinline__ Emit sh_start_prefix(const DScene& S, const RecordSet& R, ShState& s, uint32_t srcRef,
                                                uint32_t dstRef) {
    s.srcRef = srcRef;
    s.isSuffix = 0;
    s.idx = 0;
    s.count = 0;
    s.state = 0;
    s.currTraced = 0;
    s.thr = v3(1.0f);
    s.curr = packed_invalid();
    const Prefix src = load_prefix_ref(R, srcRef);
    s.srng = src.rng;
    if (!src.valid() || src.recon_idx() < 2) return emit_none();
    if (!rehydrate(S, src.primary).valid) return emit_none();
    const Inter curr = rehydrate(S, load_prefix_ref(R, dstRef).primary);
    if (!curr.valid) return emit_none();
    if (!rehydrate(S, src.recon).valid) return emit_none();
    const int rvc = (int)(src.recon_idx() - 2u);
    s.count = rvc > 0 ? (uint32_t)rvc : 0u;
    return sh_continue(S, R, s, curr);
}

__device__ __forceinline__ Emit sh_start_suffix(const DScene& S, const RecordSet& R, ShState& s, uint32_t srcRef,
                                                uint32_t dstRef) {
    s.srcRef = srcRef;
    s.isSuffix = 1;
    s.idx = 0;
    s.count = 0;
    s.state = 0;
    s.currTraced = (dstRef >> 28) == REF_PACKED_TRACED ? 1u : 0u;
    s.thr = v3(1.0f);
    s.curr = packed_invalid();
    const Inter dst = load_suffix_dst(S, R, dstRef);
    const Suffix src = load_suffix_ref(R, srcRef);
    s.srng = src.rng;
    if (!dst.valid) return emit_none();
    if (!src.valid()) return emit_none();
    if (!rehydrate(S, src.lastPrefix).valid) return emit_none();
    if (!rehydrate(S, src.recon).valid) return emit_none();
    const int rvc = (int)(src.recon_idx() - 1u);
    s.count = rvc > 0 ? (uint32_t)rvc : 0u;
    return sh_continue(S, R, s, dst);
}

__device__ __forceinline__ Emit sh_on_hit(const DScene& S, const RecordSet& R, ShState& s, const Inter& it) {
    if (!it.valid) return emit_none();
    s.currTraced = 1;
    s.idx++;
    return sh_continue(S, R, s, it);
}

// Final step: BSDFs at both ends of the reconnection and the Jacobian.  Writes f and J.
__device__ __forceinline__ void sh_on_shadow(const DScene& S, const RecordSet& R, const ShState& s, bool occluded,
                                             V3& f, float& J) {
    f = v3(0.0f);
    J = 0.0f;
    if (occluded) return;
    const Inter curr = inter_restore(S, s.curr, s.currTraced != 0);
    V3 thr = s.thr;
    if (!s.isSuffix) {
        const Prefix src = load_prefix_ref(R, s.srcRef);
        const Inter srcPrimary = rehydrate(S, src.primary);
        Inter srcRecon = rehydrate(S, src.recon);
        const V3 reconDir = normalize(srcRecon.pos - curr.pos);
        const BrdfEval b1 = ggx_eval(S, curr, reconDir);
        if (b1.pdf <= 0.0f) return;
        thr = thr * b1.brdf;
        srcRecon.dir = reconDir;
        const BrdfEval b2 = ggx_eval(S, srcRecon, src.reconOutDir);
        if (b2.pdf <= 0.0f) return;
        thr = thr * b2.brdf;
        J = recon_jacobian(srcPrimary.pos, curr.pos, srcRecon.pos, srcRecon.nrm);
        const V3 ff = thr * src.postReconF;
        f = V3{gmax(0.0f, ff.x), gmax(0.0f, ff.y), gmax(0.0f, ff.z)};
        return;
    }
    const Suffix src = load_suffix_ref(R, s.srcRef);
    const Inter srcLast = rehydrate(S, src.lastPrefix);
    Inter srcRecon = rehydrate(S, src.recon);
    const V3 reconDir = normalize(srcRecon.pos - curr.pos);
    const BrdfEval b1 = ggx_eval(S, curr, reconDir);
    if (b1.pdf <= 0.0f) return;
    thr = thr * b1.brdf;
    if (src.recon_idx() < src.length()) {
        srcRecon.dir = reconDir;
        if (is_zero(src.reconOutDir)) return;
        const BrdfEval b2 = ggx_eval(S, srcRecon, src.reconOutDir);
        if (b2.pdf <= 0.0f) return;
        thr = thr * b2.brdf;
    }
    J = recon_jacobian(srcLast.pos, curr.pos, srcRecon.pos, srcRecon.nrm);
    f = thr * src.postReconF;
}

}  // namespace cx
