// Wavefront kernels of the frame loop.
//
// Per pass: an init/setup kernel creates one job per pixel (or per pixel x neighbour x shift
// direction) and appends the first rays; then a few rounds in which the traversal kernel answers
// the queued closest-hit and shadow rays and a step kernel advances every job to its next ray or
// to its end; then a tail kernel in which every remaining job traces its rays in place until it
// ends; finally a resample kernel runs the reference's reservoir updates per pixel in the
// reference's order.  The traversal code uses a 16-entry LDS stack per thread (deeper entries
// spill to private memory), so occupancy is set by registers.
#include <algorithm>
#include <map>
#include <mutex>
#include <utility>

#include "cx_internal.h"

namespace cx {

static constexpr int WBLOCK = 128;
static_assert(WBLOCK == (int)QUEUE_BLOCK, "queue kernels and the stripe sizing agree on the block size");
static constexpr int SHORT_STACK = 16;

__device__ __forceinline__ uint32_t gid() { return blockIdx.x * blockDim.x + threadIdx.x; }

// ---- traversal of a queue ----

// The rays of wavefront w among `waves`, out of the n rays [base, base + n) of qin's index space:
// the range is cut into chunks of ch rays and wavefront w takes chunks w, w + waves, w + 2 waves,
// ...  A chunk holds the appends of one producing wavefront or a part of one, so its rays start
// close together, and the chunks of one wavefront come from all over the queue: a contiguous
// range per wavefront held rays of one region of the image (the stripes keep neighbouring
// producers together), and the wavefronts given the expensive regions finished last (1080p
// frame, measured together: 24.9 ms with 32-ray chunks, 25.6 ms with contiguous ranges).  Some
// wavefronts get one chunk more than others, so a chunk is at most a quarter of a wavefront's
// share, ch = min(TRACE_CHUNK, n / (4 waves)), at least minChunk: with 32-ray chunks throughout,
// a 512² frame (a few chunks per wavefront) took 2.10 ms instead of 1.86, and 128- or 256-ray
// chunks made the 1080p frame 25.5 and 27.9 ms; 16-64-ray chunks time alike there.  A frame that
// runs beside other views (cx_batch_frame) passes minChunk = TRACE_CHUNK: the others fill the
// CUs its late wavefronts leave, and whole chunks keep rays that start together on one
// wavefront (16 views: 129.6-131.1 ms per batch frame with 32-ray chunks against 133.5-134.2
// with a quarter share).
// The wavefront's lanes each hold one ray at a time: a lane whose ray is answered writes the
// result and idles until at least TRACE_REFILL lanes of the wavefront (or as many as rays
// remain) are idle, and then those lanes take the next rays together.  A wavefront so stays
// busy until its rays are drained, instead of waiting at every ray for its slowest lane;
// refilling in groups keeps the wavefront from stalling on a ray load at every step.
static constexpr uint32_t TRACE_REFILL = 16;
static constexpr uint32_t TRACE_CHUNK = 32;

// Wavefront w's chunks of the n rays [base, base + n): cnt rays in chunks of ch.
struct ChunkRange {
    uint32_t base, ch, cnt;
};

__device__ __forceinline__ ChunkRange chunk_range(uint32_t base, uint32_t n, uint32_t w, uint32_t waves,
                                                  uint32_t minChunk) {
    // cnt = rays of this wavefront: its chunks w + m waves < nch, the last chunk of the range
    // holding n - (nch - 1) ch rays
    ChunkRange c;
    const uint32_t share = n / (waves * 4u);
    c.base = base;
    c.ch = share < minChunk ? minChunk : (share < TRACE_CHUNK ? share : TRACE_CHUNK);
    const uint32_t nch = (n + c.ch - 1) / c.ch;
    c.cnt = 0;
    if (w < nch) {
        const uint32_t mine = (nch - 1 - w) / waves + 1;
        c.cnt = mine * c.ch;
        if (w + (mine - 1) * waves == nch - 1) c.cnt -= nch * c.ch - n;
    }
    return c;
}

// Index in qin of the wavefront's v-th ray of range c.
__device__ __forceinline__ uint32_t chunk_ray(const ChunkRange& c, uint32_t v, uint32_t w, uint32_t waves) {
    return c.base + (w + (v / c.ch) * waves) * c.ch + v % c.ch;
}

// The wavefront's closest-hit chunks and then its shadow chunks as one sequence, so a lane whose
// closest-hit ray is answered takes a shadow ray at the next refill instead of idling until its
// wavefront has finished every closest-hit ray.  A lane's ray kind is a runtime flag: the two
// kinds step through the same code, which differs only in the leaf's early exit and the result
// written.
template <int WIDTH, typename Stack>
__device__ __forceinline__ void trace_chunks(const DScene& S, const Queue& qin, const QIndex& qi, uint32_t nc,
                                             uint32_t ns, uint32_t w, uint32_t waves, uint32_t minChunk,
                                             Stack& st) {
    const ChunkRange cc = chunk_range(0u, nc, w, waves, minChunk), cs = chunk_range(nc, ns, w, waves, minChunk);
    const uint32_t cnt = cc.cnt + cs.cnt;
    const uint64_t below = (1ull << (threadIdx.x & 63)) - 1ull;
    uint32_t cursor = 0;   // wavefront-uniform, counts this wavefront's rays
    bool has = false, any = false;
    uint32_t slot = 0;
    TravRay r;
    while (true) {
        const uint64_t idle = __ballot(!has);
        if (cursor < cnt) {
            const uint32_t nIdle = (uint32_t)__popcll(idle), left = cnt - cursor;
            if (nIdle >= (left < TRACE_REFILL ? left : TRACE_REFILL)) {
                if (!has) {
                    const uint32_t v = cursor + (uint32_t)__popcll(idle & below);
                    if (v < cnt) {
                        any = v >= cc.cnt;
                        const uint32_t i = any ? chunk_ray(cs, v - cc.cnt, w, waves) : chunk_ray(cc, v, w, waves);
                        slot = in_slot(qin, qi, i);
                        const float4 a = qin.ray[slot], b = qin.ray[(size_t)qin.cap + slot];
                        trav_init<WIDTH>(S, r, V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z}, a.w, b.w);
                        has = true;
                    }
                }
                cursor += nIdle;
            }
        }
        if (!has) {
            if (cursor >= cnt) break;
            continue;
        }
        if (trav_step_rt<WIDTH>(S, r, st, any)) {
            const Hit& h = r.best;
            const bool hit = h.prim != INVALID_PRIM;
            qin.res[slot] = any ? uint4{hit ? 1u : 0u, 0u, 0u, 0u}
                                : uint4{h.prim, fb(hit ? h.t : 0.0f), fb(hit ? h.u : 0.0f), fb(hit ? h.v : 0.0f)};
            has = false;
        }
    }
}

// Answers every ray of qin in place (closest hit: prim, t, u, v; occlusion: the flag in x),
// zeroes the stripe counts of qout (the next step kernel's output queue) and adds the rays to
// the pass counters.  Each wavefront traces its chunks of the closest rays and then of the
// shadow rays in one refill loop (trace_chunks), and the grid is what fits on the device at
// once.  The 4-wide kernel is held to at least 7 waves per SIMD: 72 VGPRs, 8 bytes per lane
// more of private stack than at its natural 74 VGPRs and 6 waves (1080p frame 22.75-22.82 ms at
// 7 waves, 22.91 at 6, 23.57-23.61 at 8, 25.6 held to exactly 5).  The 2-wide kernel keeps at
// least 5.
template <int WIDTH>
__global__ __launch_bounds__(WBLOCK) __attribute__((amdgpu_waves_per_eu(WIDTH == 8 ? 1 : (WIDTH == 4 ? 7 : 5), 8))) void k_q_trace(
    DScene S, Queue qin, Queue qout, unsigned long long* counters, uint32_t minChunk) {
    __shared__ int s_st[SHORT_STACK * WBLOCK];
    __shared__ QIndex qi;
    qindex_load(qin, qi);
    const uint32_t nc = qi.pc[QSTRIPES], ns = qi.ps[QSTRIPES];
    if (blockIdx.x == 0 && threadIdx.x < QSTRIPES) {
        qout.count[threadIdx.x * QCOUNT_STRIDE] = 0u;
        qout.count[threadIdx.x * QCOUNT_STRIDE + 1] = 0u;
        if (threadIdx.x == 0 && counters) {
            if (nc) atomicAdd(&counters[0], (unsigned long long)nc);
            if (ns) atomicAdd(&counters[1], (unsigned long long)ns);
        }
    }
    TravStack<SHORT_STACK> st;
    st.lds = s_st + threadIdx.x;
    st.stride = WBLOCK;
    const uint32_t waves = gridDim.x * (WBLOCK / 64), w = blockIdx.x * (WBLOCK / 64) + (threadIdx.x >> 6);
    trace_chunks<WIDTH>(S, qin, qi, nc, ns, w, waves, minChunk, st);
}

// Grid-stride loop over the jobs of qin whose bound is uniform per block, so every lane of a
// wavefront reaches each wave-aggregated append.
#define CX_FOR_JOBS(W)                                                         \
    __shared__ QIndex _qi;                                                    \
    qindex_load((W).qin, _qi);                                                \
    const uint32_t _nc = _qi.pc[QSTRIPES], _ns = _qi.ps[QSTRIPES];            \
    const uint32_t _n = _nc + _ns;                                            \
    for (uint32_t _base = blockIdx.x * WBLOCK; _base < _n; _base += gridDim.x * WBLOCK)

__device__ __forceinline__ void emit_job(const Queue& q, uint32_t slot, uint32_t job, const Emit& e) {
    q.job[slot] = job;
    put_ray(q, slot, e);
}

__device__ __forceinline__ uint64_t gbuf_rng(const uint4& g) { return (uint64_t)g.x | ((uint64_t)g.y << 32); }

__device__ __forceinline__ void tile_xy(const FrameArgs& A, uint32_t p, uint32_t& x, uint32_t& y) {
    x = p % A.W;
    y = p / A.W;
}

// ======================================================================================
// Pass A: prefix_gen_temp_reuse (shader/prefix_gen_temp_reuse.cu)
// ======================================================================================

__device__ __forceinline__ void pg_finish(const WaveArgs& W, uint32_t p, const PgState& s) {
    PrefixRes r;
    r.s = s.pre;
    r.wSum = 0.0f;
    r.conf = 0.0f;
    store_prefix(W.prefixScratch, W.A.P, p, r);
    W.rngScratch[p] = s.rng;
}

__device__ __forceinline__ void cp_finish(const WaveArgs& W, uint32_t p, const CpState& s) {
    V3 rad = s.radiance;
    if (any_bad(rad)) rad = v3(0.0f);
    float* o = W.A.out + 3 * (size_t)p;
    V3 r = rad;
    if (W.A.lp.enableAccum) {
        const V3 old = V3{o[0], o[1], o[2]};
        const float blend = 1.0f / (float)(frame_index(W.A) + 1);
        r = blend * rad + (1.0f - blend) * old;
    }
    o[0] = r.x;
    o[1] = r.y;
    o[2] = r.z;
}

// Camera ray, then the first ray of TracePrefix (ReSTIR modes) or TraceCompletePath.
__global__ __launch_bounds__(WBLOCK) void k_w_prefix_init(WaveArgs W) {
    const FrameArgs& A = W.A;
    const uint32_t p = gid();
    const bool active = p < A.P;
    Emit e = emit_none();
    PgState pg;
    CpState cp;
    const bool restir = A.lp.rendererType != CX_PATH_TRACER;
    if (active) {
        uint32_t x, y;
        tile_xy(A, p, x, y);
        Rng rng = rng_seed(tea64(p, frame_random(A)));
        const float jx = rng.f();
        const float jy = rng.f();
        float ux = ((float)x + jx) / (float)A.W;
        float uy = ((float)y + jy) / (float)A.H;
        ux = 2.0f * ux - 1.0f;
        uy = 2.0f * uy - 1.0f;
        const FrameCam c = frame_cam(A);
        const V3 origin = c.pos;
        const V3 dir = normalize(c.U * ux + c.V * uy + c.W);
        if (restir) {
            pg_reset(pg, rng.s);
            e = pg_start(A.lp, pg, origin, dir);
            if (e.kind == EMIT_NONE) pg_finish(W, p, pg);
        } else {
            cp_reset(cp, rng.s, origin, dir);
            if (A.lp.maxPathLen > 0) e = emit_closest(origin, dir);
            else cp_finish(W, p, cp);
        }
    }
    const uint32_t slot = append(W.qout, e.kind);
    if (active && e.kind != EMIT_NONE) {
        store_fresh(W.qout, slot, restir ? pg.rng : cp.rng);
        emit_job(W.qout, slot, p, e);
    }
}

// The step of each machine on the answer r to a job's last ray (origin o, direction d): the
// reference's logic up to the job's next trace call or its end, where it stores the result.
// Shared by the step kernels and the tail kernel.
__device__ __forceinline__ Emit pg_advance(const WaveArgs& W, uint32_t job, PgState& s, uint4 r, V3 o, V3 d) {
    const Emit e = s.state == PG_TRACE ? pg_on_hit(W.A.S, W.A.lp, s, inter_from_hit(W.A.S, r, d))
                                       : pg_on_shadow(W.A.S, W.A.lp, s, r.x != 0);
    if (e.kind == EMIT_NONE) pg_finish(W, job, s);
    return e;
}

__global__ __launch_bounds__(WBLOCK) void k_w_pg_step(WaveArgs W) {
    CX_FOR_JOBS(W) {
    const uint32_t i = _base + threadIdx.x;
    const bool active = i < _n;
    Emit e = emit_none();
    PgState s;
    uint32_t job = 0;
    if (active) {
        const uint32_t slot = in_slot(W.qin, _qi, i);
        job = W.qin.job[slot];
        s = pg_load(W.qin, slot, W.A.lp);
        V3 o, d;
        get_ray(W.qin, slot, o, d);
        e = pg_advance(W, job, s, W.qin.res[slot], o, d);
    }
    const uint32_t slot = append(W.qout, e.kind);
    if (active && e.kind != EMIT_NONE) {
        store_ctx(W.qout, slot, pg_pack(s));
        emit_job(W.qout, slot, job, e);
    }
    }
}

__device__ __forceinline__ Emit cp_advance(const WaveArgs& W, uint32_t job, CpState& s, uint4 r, V3 o, V3 d) {
    const Emit e = s.state == CP_TRACE ? cp_on_hit(W.A.S, W.A.lp, s, inter_from_hit(W.A.S, r, d))
                                       : cp_on_shadow(W.A.S, s, r.x != 0);
    if (e.kind == EMIT_NONE) cp_finish(W, job, s);
    return e;
}

__global__ __launch_bounds__(WBLOCK) void k_w_cp_step(WaveArgs W) {
    CX_FOR_JOBS(W) {
    const uint32_t i = _base + threadIdx.x;
    const bool active = i < _n;
    Emit e = emit_none();
    CpState s;
    uint32_t job = 0;
    if (active) {
        const uint32_t slot = in_slot(W.qin, _qi, i);
        job = W.qin.job[slot];
        s = cp_load(W.qin, slot);
        V3 o, d;
        get_ray(W.qin, slot, o, d);
        e = cp_advance(W, job, s, W.qin.res[slot], o, d);
    }
    const uint32_t slot = append(W.qout, e.kind);
    if (active && e.kind != EMIT_NONE) {
        store_ctx(W.qout, slot, cp_pack(s));
        emit_job(W.qout, slot, job, e);
    }
    }
}

// Temporal decision of PrefixGenTempReuse (prefix_gen_temp_reuse.cu:31-54).
__device__ __forceinline__ bool prefix_temporal(const WaveArgs& W, uint32_t p, const PrefixRes& prev) {
    const cx_launch_params& lp = W.A.lp;
    if (!lp.prefixEnableTemporal) return false;
    const bool skipNee = is_cond(lp) && prev.s.nee();
    return !(prev.wSum <= 0.0f || !prev.s.valid() || skipNee);
}

__device__ __forceinline__ void shift_finish(const WaveArgs& W, uint32_t job, V3 f, float J) {
    W.shiftRes[job] = float4{f.x, f.y, f.z, J};
}

// Both temporal shifts of a pixel (prefix_gen_temp_reuse.cu:58-61).
__global__ __launch_bounds__(WBLOCK) void k_w_prefix_temporal_setup(WaveArgs W) {
    const uint32_t p = gid();
    const bool px = p < W.A.P;
    bool temporal = false;
    if (px) {
        const uint32_t cf = load_prefix_flags(W.prefixScratch, W.A.P, p);
        if ((cf >> 16) & 1u) temporal = prefix_temporal(W, p, load_prefix(W.A.prefixRes[W.A.back], W.A.P, p));
        W.pixFlags[p] = temporal ? 1u : 0u;
    }
    for (uint32_t d = 0; d < 2; d++) {
        Emit e = emit_none();
        ShState s;
        const uint32_t job = 2 * p + d;
        if (temporal) {
            const uint32_t back = make_ref(REF_PREFIX_BACK, p), canon = make_ref(REF_PREFIX_SCRATCH, p);
            e = d == 0 ? sh_start_prefix(W.A.S, W.R, s, back, canon) : sh_start_prefix(W.A.S, W.R, s, canon, back);
            if (e.kind == EMIT_NONE) shift_finish(W, job, v3(0.0f), 0.0f);
        }
        const uint32_t slot = append(W.qout, e.kind);
        if (temporal && e.kind != EMIT_NONE) {
            store_ctx(W.qout, slot, sh_pack(s));
            emit_job(W.qout, slot, job, e);
        }
    }
}

__device__ __forceinline__ Emit sh_advance(const WaveArgs& W, uint32_t job, ShState& s, uint4 r, V3 o, V3 d) {
    if (s.state == SH_TRACE) {
        const Emit e = sh_on_hit(W.A.S, W.R, s, inter_from_hit(W.A.S, r, d));
        if (e.kind == EMIT_NONE) shift_finish(W, job, v3(0.0f), 0.0f);
        return e;
    }
    V3 f;
    float J;
    sh_on_shadow(W.A.S, W.R, s, r.x != 0, f, J);
    shift_finish(W, job, f, J);
    return emit_none();
}

__global__ __launch_bounds__(WBLOCK) void k_w_sh_step(WaveArgs W) {
    CX_FOR_JOBS(W) {
    const uint32_t i = _base + threadIdx.x;
    const bool active = i < _n;
    Emit e = emit_none();
    ShState s;
    uint32_t job = 0;
    if (active) {
        const uint32_t slot = in_slot(W.qin, _qi, i);
        job = W.qin.job[slot];
        s = sh_unpack(load_ctx<SH_CHUNKS>(W.qin, slot));
        V3 o, d;
        get_ray(W.qin, slot, o, d);
        e = sh_advance(W, job, s, W.qin.res[slot], o, d);
    }
    const uint32_t slot = append(W.qout, e.kind);
    if (active && e.kind != EMIT_NONE) {
        store_ctx(W.qout, slot, sh_pack(s));
        emit_job(W.qout, slot, job, e);
    }
    }
}

__device__ __forceinline__ void write_out(const FrameArgs& A, uint32_t p, V3 rad) {
    float* o = A.out + 3 * (size_t)p;
    V3 r = rad;
    if (A.lp.enableAccum) {
        const V3 old = V3{o[0], o[1], o[2]};
        const float blend = 1.0f / (float)(frame_index(A) + 1);
        r = blend * rad + (1.0f - blend) * old;
    }
    o[0] = r.x;
    o[1] = r.y;
    o[2] = r.z;
}

// Reservoir updates of PrefixGenTempReuse and the raygen epilogue (prefix_gen_temp_reuse.cu:24-172).
__global__ __launch_bounds__(WBLOCK) void k_w_prefix_resample(WaveArgs W) {
    const FrameArgs& A = W.A;
    const uint32_t p = gid();
    if (p >= A.P) return;
    const cx_launch_params& lp = A.lp;
    const PrefixRes cres = load_prefix(W.prefixScratch, A.P, p);
    const Prefix& canon = cres.s;
    Rng rng;
    rng.s = W.rngScratch[p];
    const uint32_t oldFlags = A.gbuf[p].w;
    uint32_t flags = oldFlags & CANON_SUFFIX_VALID;
    if (canon.valid()) flags |= CANON_PREFIX_VALID;
    if ((oldFlags & CANON_PRIMARY_VALID) || canon.length() >= 1) flags |= CANON_PRIMARY_VALID;
    uint4* front = A.prefixRes[A.front];
    uint4* back = A.prefixRes[A.back];
    bool validPrimary = false;
    V3 rad = v3(0.0f);
    if (canon.valid()) {
        const float canonPHat = luminance(canon.f);
        PrefixRes curr = load_prefix(front, A.P, p);
        curr.wSum = 0.0f;
        curr.conf = 0.0f;
        if (!W.pixFlags[p]) {
            res_update(curr, canon, canonPHat / canon.p, rng, 1.0f);
            res_finalize(curr);
        } else {
            const PrefixRes prev = load_prefix(back, A.P, p);
            const float4 a = W.shiftRes[2 * p], b = W.shiftRes[2 * p + 1];
            const V3 fFromCanonOfPrev = V3{a.x, a.y, a.z};
            const float jPrevToCanon = a.w;
            const V3 fFromPrevOfCanon = V3{b.x, b.y, b.z};
            const float jCanonToPrev = b.w;
            const float pFromCanonOfPrev = luminance(fFromCanonOfPrev) * jPrevToCanon;
            const float K = 1.0f;
            const float prevMis = neighbor_mis(fFromCanonOfPrev, prev.s.f, jPrevToCanon, K, 1.0f, prev.conf);
            const float canonMis = 1.0f + canonical_mis(canon.f, fFromPrevOfCanon, jCanonToPrev, K, 1.0f, prev.conf);
            const float canonRis = canonMis * canonPHat / canon.p;
            res_update(curr, canon, canonRis, rng, 1.0f);
            const float prevRis = prevMis * pFromCanonOfPrev * prev.wSum;
            Prefix shifted = prev.s;
            shifted.f = fFromCanonOfPrev;
            shifted.primary = canon.primary;
            res_update(curr, shifted, prevRis, rng, prev.conf);
            if (curr.wSum > 0.0f) {
                curr.wSum /= K + 1.0f;
                res_finalize(curr);
            }
        }
        store_prefix(front, A.P, p, curr);
        if (A.copyPrefixToBack) store_prefix(back, A.P, p, curr);
        validPrimary = true;
        rad = curr.wSum * curr.s.f;
    } else if (A.copyPrefixToBack) {
        store_prefix(back, A.P, p, load_prefix(front, A.P, p));
    }
    if (is_pt(lp) && !lp.prefixEnableSpatial) write_out(A, p, validPrimary ? rad : v3(0.0f));
    A.gbuf[p] = uint4{(uint32_t)rng.s, (uint32_t)(rng.s >> 32), validPrimary ? 1u : 0u, flags};
}

// ======================================================================================
// Pass B / D: spatial reuse (prefix_spatial_reuse.cu, suffix_spatial_reuse.cu)
// ======================================================================================

// Neighbour selection and validity loop; one shift job pair per valid neighbour.
template <bool SUFFIX>
__global__ __launch_bounds__(WBLOCK) void k_w_spatial_setup(WaveArgs W) {
    const FrameArgs& A = W.A;
    const cx_launch_params& lp = A.lp;
    const uint32_t p = gid();
    const bool px = p < A.P;
    bool ok = false;
    uint32_t validFlags = 0;
    const uint32_t S = W.S;
    if (px) {
        const uint4 g = A.gbuf[p];
        if (!SUFFIX) {
            ok = ((load_prefix_flags(A.prefixRes[A.front], A.P, p) >> 16) & 1u) && (g.w & CANON_PREFIX_VALID) &&
                 (g.w & CANON_PRIMARY_VALID);
        } else {
            const uint32_t pf = load_prefix_flags(A.prefixRes[A.front], A.P, p);
            ok = ((pf >> 16) & 1u) && !((pf >> 17) & 1u) &&
                 ((load_suffix_flags(A.suffixRes[A.front], A.P, p) >> 16) & 1u) && (g.w & CANON_SUFFIX_VALID);
        }
        if (ok) {
            Rng rng;
            rng.s = gbuf_rng(g);
            uint32_t x, y;
            tile_xy(A, p, x, y);
            for (uint32_t i = 0; i < S; ++i) {
                uint32_t nx, ny;
                select_neighbor(x, y, lp.windowRadius, rng, nx, ny);
                uint32_t np = INVALID_PRIM;
                if (nx < A.W && ny < A.H) {
                    const uint32_t q = ny * A.W + nx;
                    if (A.gbuf[q].z) {
                        if (!SUFFIX) {
                            const uint32_t nf = load_prefix_flags(A.prefixRes[A.back], A.P, q);
                            if (((nf >> 16) & 1u) && !(is_cond(lp) && ((nf >> 17) & 1u))) np = q;
                        } else {
                            if ((load_suffix_flags(A.suffixRes[A.back], A.P, q) >> 16) & 1u) np = q;
                        }
                    }
                }
                if (np != INVALID_PRIM) validFlags |= 1u << i;
                W.nbr[(size_t)p * S + i] = np;
            }
            W.rngScratch[p] = rng.s;
        }
        W.pixFlags[p] = ok ? (1u | (validFlags << 1)) : 0u;
    }
    for (uint32_t i = 0; i < S; ++i) {
        for (uint32_t d = 0; d < 2; d++) {
            Emit e = emit_none();
            ShState s;
            const uint32_t job = ((uint32_t)p * S + i) * 2 + d;
            const bool want = ok && ((validFlags >> i) & 1u);
            if (want) {
                const uint32_t np = W.nbr[(size_t)p * S + i];
                if (!SUFFIX) {
                    const uint32_t neigh = make_ref(REF_PREFIX_BACK, np), curr = make_ref(REF_PREFIX_FRONT, p);
                    e = d == 0 ? sh_start_prefix(A.S, W.R, s, neigh, curr) : sh_start_prefix(A.S, W.R, s, curr, neigh);
                } else {
                    const uint32_t neigh = make_ref(REF_SUFFIX_BACK, np), curr = make_ref(REF_SUFFIX_FRONT, p);
                    e = d == 0 ? sh_start_suffix(A.S, W.R, s, neigh, curr) : sh_start_suffix(A.S, W.R, s, curr, neigh);
                }
                if (e.kind == EMIT_NONE) shift_finish(W, job, v3(0.0f), 0.0f);
            }
            const uint32_t slot = append(W.qout, e.kind);
            if (want && e.kind != EMIT_NONE) {
                store_ctx(W.qout, slot, sh_pack(s));
                emit_job(W.qout, slot, job, e);
            }
        }
    }
}

__global__ __launch_bounds__(WBLOCK) void k_w_prefix_spatial_resample(WaveArgs W) {
    const FrameArgs& A = W.A;
    const cx_launch_params& lp = A.lp;
    const uint32_t p = gid();
    if (p >= A.P) return;
    const uint32_t pf = W.pixFlags[p];
    V3 rad = v3(0.0f);
    bool ok = false;
    if (pf & 1u) {
        uint4* front = A.prefixRes[A.front];
        const uint4* back = A.prefixRes[A.back];
        PrefixRes curr = load_prefix(front, A.P, p);
        const Prefix currPrefix = curr.s;
        Rng rng;
        rng.s = W.rngScratch[p];
        const float currWSum = curr.wSum;
        const float currConf = curr.conf;
        curr.wSum = 0.0f;
        curr.conf = 0.0f;
        const uint32_t validFlags = pf >> 1;
        const float K = (float)__popc(validFlags);
        float canonMis = 1.0f;
        for (uint32_t i = 0; i < W.S; ++i) {
            if (!((validFlags >> i) & 1u)) continue;
            const uint32_t np = W.nbr[(size_t)p * W.S + i];
            const PrefixRes nres = load_prefix(back, A.P, np);
            const float4 a = W.shiftRes[((size_t)p * W.S + i) * 2], b = W.shiftRes[((size_t)p * W.S + i) * 2 + 1];
            const V3 fFromCanonOfNeigh = V3{a.x, a.y, a.z};
            const float jNC = a.w;
            const V3 fFromNeighOfCanon = V3{b.x, b.y, b.z};
            const float jCN = b.w;
            const float pFromCanonOfNeigh = luminance(fFromCanonOfNeigh) * jNC;
            float nMis = neighbor_mis(fFromCanonOfNeigh, nres.s.f, jNC, K, currConf, nres.conf);
            if (bad(nMis)) nMis = 0.0f;
            const float nRis = nMis * pFromCanonOfNeigh * nres.wSum;
            Prefix shifted = nres.s;
            shifted.f = fFromCanonOfNeigh;
            shifted.primary = currPrefix.primary;
            res_update(curr, shifted, nRis, rng, nres.conf);
            canonMis += canonical_mis(currPrefix.f, fFromNeighOfCanon, jCN, K, currConf, nres.conf);
        }
        if (bad(canonMis)) canonMis = 0.0f;
        const float canonRis = canonMis * currWSum * luminance(currPrefix.f);
        res_update(curr, currPrefix, canonRis, rng, currConf);
        if (curr.wSum > 0.0f) {
            curr.wSum /= K + 1.0f;
            res_finalize(curr);
        }
        store_prefix(front, A.P, p, curr);
        const uint4 g = A.gbuf[p];
        A.gbuf[p] = uint4{(uint32_t)rng.s, (uint32_t)(rng.s >> 32), g.z, g.w};
        ok = true;
        rad = curr.wSum * curr.s.f;
    }
    if (is_pt(lp) && lp.spatialRoundIdx == lp.prefixSpatialRounds - 1) write_out(A, p, ok ? rad : v3(0.0f));
}

__device__ __forceinline__ void invalidate_res(SuffixRes& r) {   // suffix_gen_temp_reuse.cu:12-17
    r.wSum = 0.0f;
    r.conf = 0.0f;
    r.s.set_valid(false);
}

__global__ __launch_bounds__(WBLOCK) void k_w_suffix_spatial_resample(WaveArgs W) {
    const FrameArgs& A = W.A;
    const uint32_t p = gid();
    if (p >= A.P) return;
    const uint32_t pf = W.pixFlags[p];
    uint4* front = A.suffixRes[A.front];
    const uint4* back = A.suffixRes[A.back];
    SuffixRes curr = load_suffix(front, A.P, p);
    if (!(pf & 1u)) {
        invalidate_res(curr);
        store_suffix(front, A.P, p, curr);
        return;
    }
    const Suffix currSuffix = curr.s;
    Rng rng;
    rng.s = W.rngScratch[p];
    const float currWSum = curr.wSum;
    const float currConf = curr.conf;
    curr.wSum = 0.0f;
    curr.conf = 0.0f;
    const uint32_t validFlags = pf >> 1;
    const float K = (float)__popc(validFlags);
    float canonMis = 1.0f;
    for (uint32_t i = 0; i < W.S; ++i) {
        if (!((validFlags >> i) & 1u)) continue;
        const uint32_t np = W.nbr[(size_t)p * W.S + i];
        const SuffixRes nres = load_suffix(back, A.P, np);
        const float4 a = W.shiftRes[((size_t)p * W.S + i) * 2], b = W.shiftRes[((size_t)p * W.S + i) * 2 + 1];
        const V3 fFromCanonOfNeigh = V3{a.x, a.y, a.z};
        const float jNC = a.w;
        const V3 fFromNeighOfCanon = V3{b.x, b.y, b.z};
        const float jCN = b.w;
        const float pFromCanonOfNeigh = luminance(fFromCanonOfNeigh) * jNC;
        float nMis = neighbor_mis(fFromCanonOfNeigh, nres.s.f, jNC, K, currConf, nres.conf);
        if (bad(nMis)) nMis = 0.0f;
        const float nRis = nMis * pFromCanonOfNeigh * nres.wSum;
        Suffix shifted = nres.s;
        shifted.lastPrefix = currSuffix.lastPrefix;
        shifted.f = fFromCanonOfNeigh;
        res_update(curr, shifted, nRis, rng, nres.conf);
        canonMis += canonical_mis(currSuffix.f, fFromNeighOfCanon, jCN, K, currConf, nres.conf);
    }
    if (bad(canonMis)) canonMis = 0.0f;
    const float canonRis = canonMis * currWSum * luminance(currSuffix.f);
    res_update(curr, currSuffix, canonRis, rng, currConf);
    if (curr.wSum > 0.0f) {
        curr.wSum /= K + 1.0f;
        res_finalize(curr);
    }
    store_suffix(front, A.P, p, curr);
    const uint4 g = A.gbuf[p];
    A.gbuf[p] = uint4{(uint32_t)rng.s, (uint32_t)(rng.s >> 32), g.z, g.w};
}

// ======================================================================================
// Pass C: suffix_gen_temp_reuse (shader/suffix_gen_temp_reuse.cu)
// ======================================================================================

__device__ __forceinline__ void sg_finish_store(const WaveArgs& W, uint32_t p, const SgState& s) {
    SuffixRes r;
    r.s = s.suf;
    r.wSum = 0.0f;
    r.conf = 0.0f;
    store_suffix(W.suffixScratch, W.A.P, p, r);
    W.rngScratch[p] = s.rng;
}

__device__ __forceinline__ void sg_init_state(SgState& s, uint64_t rng) {
    s.rng = rng;
    s.suf.recon = packed_invalid();
    s.suf.reconOutDir = v3(0.0f);
    s.it = packed_invalid();
    s.lightDir = v3(0.0f);
    s.lightLen = 0.0f;
    s.color = v3(0.0f);
    s.brdf = v3(0.0f);
    s.maxLen = 0;
}

// TraceSuffix from the front prefix of every pixel whose prefix is valid and not NEE-terminated.
// gather != 0: from the canonical gather prefix in the prefix scratch instead.
__global__ __launch_bounds__(WBLOCK) void k_w_suffix_init(WaveArgs W, int gather) {
    const FrameArgs& A = W.A;
    const uint32_t p = gid();
    const bool px = p < A.P;
    Emit e = emit_none();
    SgState s;
    bool run = false;
    if (px) {
        if (!gather) {
            const uint32_t pf = load_prefix_flags(A.prefixRes[A.front], A.P, p);
            run = ((pf >> 16) & 1u) && !((pf >> 17) & 1u);
            W.pixFlags[p] = run ? 1u : 0u;
            if (run) {
                sg_init_state(s, gbuf_rng(A.gbuf[p]));
                e = sg_start(A.S, A.lp, s, load_prefix(A.prefixRes[A.front], A.P, p).s);
            }
        } else {
            const uint32_t pf = load_prefix_flags(W.prefixScratch, A.P, p);
            run = (W.pixFlags[p] & 1u) && ((pf >> 16) & 1u) && !((pf >> 17) & 1u);
            if (run) {
                sg_init_state(s, W.rngScratch[p]);
                e = sg_start(A.S, A.lp, s, load_prefix(W.prefixScratch, A.P, p).s);
            }
        }
        if (run && e.kind == EMIT_NONE) sg_finish_store(W, p, s);
    }
    const uint32_t slot = append(W.qout, e.kind);
    if (run && e.kind != EMIT_NONE) {
        store_ctx(W.qout, slot, sg_pack(s));
        emit_job(W.qout, slot, p, e);
    }
}

__device__ __forceinline__ Emit sg_advance(const WaveArgs& W, uint32_t job, SgState& s, uint4 r, V3 o, V3 d) {
    Emit e;
    if (s.state == SG_TRACE) e = sg_on_hit(W.A.S, W.A.lp, s, inter_from_hit(W.A.S, r, d), o);
    else if (s.state == SG_NEE) e = sg_on_shadow(W.A.S, s, r.x != 0);
    else e = sg_on_recon(s, inter_from_hit(W.A.S, r, d));
    if (e.kind == EMIT_NONE) sg_finish_store(W, job, s);
    return e;
}

__global__ __launch_bounds__(WBLOCK) void k_w_sg_step(WaveArgs W) {
    CX_FOR_JOBS(W) {
    const uint32_t i = _base + threadIdx.x;
    const bool active = i < _n;
    Emit e = emit_none();
    SgState s;
    uint32_t job = 0;
    if (active) {
        const uint32_t slot = in_slot(W.qin, _qi, i);
        job = W.qin.job[slot];
        s = sg_unpack(load_ctx<SG_CHUNKS>(W.qin, slot), W.qin.ray[slot], W.qin.ray[(size_t)W.qin.cap + slot]);
        V3 o, d;
        get_ray(W.qin, slot, o, d);
        e = sg_advance(W, job, s, W.qin.res[slot], o, d);
    }
    const uint32_t slot = append(W.qout, e.kind);
    if (active && e.kind != EMIT_NONE) {
        store_ctx(W.qout, slot, sg_pack(s));
        emit_job(W.qout, slot, job, e);
    }
    }
}

__global__ __launch_bounds__(WBLOCK) void k_w_suffix_temporal_setup(WaveArgs W) {
    const FrameArgs& A = W.A;
    const uint32_t p = gid();
    const bool px = p < A.P;
    bool temporal = false;
    if (px) {
        const bool ran = W.pixFlags[p] & 1u;
        if (ran && ((load_suffix_flags(W.suffixScratch, A.P, p) >> 16) & 1u) && A.lp.suffixEnableTemporal) {
            const SuffixRes prev = load_suffix(A.suffixRes[A.back], A.P, p);
            temporal = !(prev.wSum <= 0.0f || !prev.s.valid());
        }
        W.pixFlags[p] = (ran ? 1u : 0u) | (temporal ? 2u : 0u);
    }
    for (uint32_t d = 0; d < 2; d++) {
        Emit e = emit_none();
        ShState s;
        const uint32_t job = 2 * p + d;
        if (temporal) {
            const uint32_t back = make_ref(REF_SUFFIX_BACK, p), canon = make_ref(REF_SUFFIX_SCRATCH, p);
            e = d == 0 ? sh_start_suffix(A.S, W.R, s, back, canon) : sh_start_suffix(A.S, W.R, s, canon, back);
            if (e.kind == EMIT_NONE) shift_finish(W, job, v3(0.0f), 0.0f);
        }
        const uint32_t slot = append(W.qout, e.kind);
        if (temporal && e.kind != EMIT_NONE) {
            store_ctx(W.qout, slot, sh_pack(s));
            emit_job(W.qout, slot, job, e);
        }
    }
}

__global__ __launch_bounds__(WBLOCK) void k_w_suffix_resample(WaveArgs W) {
    const FrameArgs& A = W.A;
    const uint32_t p = gid();
    if (p >= A.P) return;
    const uint4 g = A.gbuf[p];
    uint32_t flags = g.w;
    Rng rng;
    rng.s = gbuf_rng(g);
    uint4* front = A.suffixRes[A.front];
    uint4* back = A.suffixRes[A.back];
    SuffixRes curr = load_suffix(front, A.P, p);
    const uint32_t pf = W.pixFlags[p];
    if (!(pf & 1u)) {
        invalidate_res(curr);
    } else {
        const SuffixRes cres = load_suffix(W.suffixScratch, A.P, p);
        const Suffix& canon = cres.s;
        rng.s = W.rngScratch[p];
        flags = (flags & ~CANON_SUFFIX_VALID) | (canon.valid() ? CANON_SUFFIX_VALID : 0u);
        if (!canon.valid()) {
            invalidate_res(curr);
        } else {
            const float canonPHat = luminance(canon.f);
            curr.wSum = 0.0f;
            curr.conf = 0.0f;
            if (!(pf & 2u)) {
                res_update(curr, canon, canonPHat / canon.p, rng, 1.0f);
                res_finalize(curr);
            } else {
                const SuffixRes prev = load_suffix(back, A.P, p);
                const float4 a = W.shiftRes[2 * p], b = W.shiftRes[2 * p + 1];
                const V3 fFromCanonOfPrev = V3{a.x, a.y, a.z};
                const float jPrevToCanon = a.w;
                const V3 fFromPrevOfCanon = V3{b.x, b.y, b.z};
                const float jCanonToPrev = b.w;
                const float pFromCanonOfPrev = luminance(fFromCanonOfPrev) * jPrevToCanon;
                const float K = 1.0f;
                const float prevMis = neighbor_mis(fFromCanonOfPrev, prev.s.f, jPrevToCanon, K, 1.0f, prev.conf);
                const float canonMis =
                    1.0f + canonical_mis(canon.f, fFromPrevOfCanon, jCanonToPrev, K, 1.0f, prev.conf);
                const float canonRis = canonMis * canonPHat / canon.p;
                res_update(curr, canon, canonRis, rng, 1.0f);
                const float prevRis = prevMis * pFromCanonOfPrev * prev.wSum;
                Suffix shifted = prev.s;
                shifted.lastPrefix = canon.lastPrefix;
                shifted.f = fFromCanonOfPrev;
                res_update(curr, shifted, prevRis, rng, prev.conf);
                if (curr.wSum > 0.0f) {
                    curr.wSum /= K + 1.0f;
                    res_finalize(curr);
                }
            }
        }
    }
    store_suffix(front, A.P, p, curr);
    store_suffix(back, A.P, p, curr);
    A.gbuf[p] = uint4{(uint32_t)rng.s, (uint32_t)(rng.s >> 32), g.z, flags};
}

// ======================================================================================
// Pass F: final_gather (shader/final_gather.cu)
// ======================================================================================

// GetRadiance prologue (final_gather.cu:290-308) and the first ray of the canonical prefix.
__global__ __launch_bounds__(WBLOCK) void k_w_gather_init(WaveArgs W) {
    const FrameArgs& A = W.A;
    const uint32_t p = gid();
    const bool px = p < A.P;
    Emit e = emit_none();
    PgState pg;
    bool run = false;
    if (px) {
        const uint4 g = A.gbuf[p];
        run = (g.w & CANON_PRIMARY_VALID) != 0;
        W.pixFlags[p] = run ? 1u : 0u;
        W.radScratch[p] = float4{0.0f, 0.0f, 0.0f, 0.0f};
        if (run) {
            uint32_t x, y;
            tile_xy(A, p, x, y);
            Rng rng;
            rng.s = gbuf_rng(g);
            const float jx = rng.f();
            const float jy = rng.f();
            float ux = ((float)x + jx) / (float)A.W;
            float uy = ((float)y + jy) / (float)A.H;
            ux = 2.0f * ux - 1.0f;
            uy = 2.0f * uy - 1.0f;
            const FrameCam c = frame_cam(A);
            const V3 dir = normalize(c.U * ux + c.V * uy + c.W);
            W.camDir[p] = float4{dir.x, dir.y, dir.z, 0.0f};
            pg_reset(pg, rng.s);
            e = pg_start(A.lp, pg, c.pos, dir);
            if (e.kind == EMIT_NONE) pg_finish(W, p, pg);
        }
    }
    const uint32_t slot = append(W.qout, e.kind);
    if (run && e.kind != EMIT_NONE) {
        store_fresh(W.qout, slot, pg.rng);
        emit_job(W.qout, slot, p, e);
    }
}

// Deferred siblings of the gather's entry query kept in LDS per thread; deeper ones (rare: the
// walk defers one per level only where both children stay in reach) go to private memory.
static constexpr int GATHER_QUERY_SHORT = 8;

template <int KR>
__device__ __forceinline__ uint32_t query_neighbors(const WaveArgs& W, uint32_t p, V3 q, int* stack, int stride) {
    const uint32_t k = W.A.k;
    uint32_t* nIdx = W.A.neighIdx + (size_t)p * k;
    float* nDist = W.A.neighDist + (size_t)p * k;
    QueryStack<GATHER_QUERY_SHORT> st;
    st.lds = stack;
    st.stride = stride;
    const uint32_t count = entries_query<KR>(W.A.tree, q, k, st, nIdx, nDist);
    W.A.neighCount[p] = count;
    return count;
}

// neighbor_stat over the queries of a whole wavefront (every lane calls it; `active` marks the
// lanes that queried): the sum, minimum and maximum are reduced across the lanes and added by
// lane 0 with one atomic each, instead of three atomics per query on 64 striped counter sets.
__device__ __forceinline__ void neighbor_stats_wave(const WaveArgs& W, bool active, uint32_t count) {
    if (!W.A.lp.trackPrefixStats) return;
    unsigned long long sum = active ? count : 0ull;
    uint32_t mn = active ? count : 0xFFFFFFFFu, mx = active ? count : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        sum += __shfl_xor(sum, off, 64);
        mn = min(mn, (uint32_t)__shfl_xor((int)mn, off, 64));
        mx = max(mx, (uint32_t)__shfl_xor((int)mx, off, 64));
    }
    if (__ballot(active) != 0ull && (threadIdx.x & 63u) == 0u) {
        unsigned long long* c =
            W.A.counters + (((blockIdx.x * blockDim.x + threadIdx.x) >> 6) % COUNTER_STRIPES) * NUM_COUNTERS;
        atomicAdd(&c[STAT_TOTAL], sum);
        atomicMin(&c[STAT_MIN], (unsigned long long)mn);
        atomicMax(&c[STAT_MAX], (unsigned long long)mx);
    }
}

// Record-major copy of the front suffix reservoirs (the final gather's neighbours, read at
// random: one 128-byte line per record instead of one line per chunk).  One thread per 16-byte
// word of the copy, so a wavefront writes 1 KiB of whole lines and reads 8 runs of 8 records
// from the chunk arrays.  The final gather reads the neighbours through it and writes no suffix
// reservoir, so the copy stays exact.
__global__ __launch_bounds__(WBLOCK) void k_w_suffix_aos(WaveArgs W) {
    const FrameArgs& A = W.A;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)A.P * SUFFIX_AOS_STRIDE) return;
    const uint32_t p = (uint32_t)(i / SUFFIX_AOS_STRIDE), c = (uint32_t)(i % SUFFIX_AOS_STRIDE);
    W.suffixAos[i] = c < (uint32_t)SUFFIX_CHUNKS ? ld4(A.suffixRes[A.front], A.P, (int)c, p) : uint4{0u, 0u, 0u, 0u};
}

// Whether pixel p queries neighbours in this gather: gIdx 0, a valid canonical prefix that did
// not end in NEE and a valid canonical suffix (final_gather.cu:90-133); gIdx > 0, a valid gather
// prefix (final_gather.cu:209-246).
__device__ __forceinline__ bool gather_runs(const WaveArgs& W, uint32_t p, uint32_t pf) {
    const FrameArgs& A = W.A;
    if (!(pf & 1u) || A.k == 0) return false;
    if (W.gIdx == 0) {
        const uint32_t f = load_prefix_flags(W.prefixScratch, A.P, p);
        return ((f >> 16) & 1u) && !((f >> 17) & 1u) && ((load_suffix_flags(W.suffixScratch, A.P, p) >> 16) & 1u);
    }
    return (pf & 4u) != 0u;
}

// The point pixel p queries around: the canonical prefix's last vertex (gIdx 0) or the gather
// prefix's traced last vertex (gIdx > 0).
__device__ __forceinline__ V3 gather_query_point(const WaveArgs& W, uint32_t p) {
    const FrameArgs& A = W.A;
    if (W.gIdx == 0) return rehydrate(A.S, load_suffix(W.suffixScratch, A.P, p).s.lastPrefix).pos;
    const uint4 a = W.packedScratch[2 * (size_t)p], b = W.packedScratch[2 * (size_t)p + 1];
    const Packed pk = Packed{a.x, bf(a.y), bf(a.z), V3{bf(a.w), bf(b.x), bf(b.y)}};
    return inter_restore(A.S, pk, true).pos;
}

// The querying pixels (a third of the pixels on the 1080p workload), so that the query and the
// shift starts run on them alone: each is marked (pixFlags bit 3, "queried"; validMask 0 for the
// others) and keyed by the Morton code of its query point, the others get the key 0xFFFFFFFF,
// and the entry sort's radix sort (launch_key_sort) drops those and orders the rest by (key,
// pixel) into gatherList.  The query points sit one bounce past the primary hits, all over the
// scene for the pixels of one wavefront; in Morton order a wavefront's tree walks and its
// neighbours' records share cache lines, and so do the shifts they start (1080p frame: query
// 0.74 -> 0.55 ms, shift steps 1.72 -> 1.60 ms, for 0.11 ms of sorting).
__global__ __launch_bounds__(WBLOCK) void k_w_gather_key(WaveArgs W) {
    const uint32_t p = gid();
    if (p >= W.A.P) return;
    bool run = false;
    const uint32_t pf = W.pixFlags[p];
    if (pf & 1u) {
        run = gather_runs(W, p, pf);
        W.pixFlags[p] = (pf & 7u) | (run ? 8u : 0u);
        if (!run) W.validMask[p] = 0u;
    }
    uint32_t key = 0xFFFFFFFFu;
    if (run) {
        const V3 q = gather_query_point(W, p);
        W.gatherPos[p] = float4{q.x, q.y, q.z, 0.0f};
        key = morton30(q, W.mortonLo, W.mortonScale);
    }
    W.gatherKeys[p] = key;
}

__device__ __forceinline__ uint32_t gather_total(const WaveArgs& W) { return *W.gatherCount; }

// The neighbour query of each listed pixel and its valid-neighbour mask (final_gather.cu:90-133;
// gIdx > 0: final_gather.cu:209-246), one thread per listed pixel.  The tree walk is a chain of
// dependent fetches, so it runs apart from the register-heavy shift starts (k_w_gather_setup), at
// more waves per SIMD; its deferred siblings take GATHER_QUERY_SHORT stack entries of LDS per
// thread (dynamic shared memory, gather_query_lds bytes).
template <int KR>
__global__ __launch_bounds__(WBLOCK) __attribute__((amdgpu_waves_per_eu(KR <= 8 ? 8 : 1, 8))) void k_w_gather_query(
    WaveArgs W) {
    extern __shared__ int s_st[];
    const FrameArgs& A = W.A;
    const uint32_t j = gid();
    const bool active = j < gather_total(W);
    uint32_t count = 0;
    if (active) {
        const uint32_t p = W.gatherList[j], k = A.k;
        const float4 qp = W.gatherPos[p];
        const V3 q = V3{qp.x, qp.y, qp.z};
        count = query_neighbors<KR>(W, p, q, s_st + threadIdx.x, WBLOCK);
        uint32_t validFlags = 0;
        for (uint32_t i = 0; i < count; ++i)
            if ((load_suffix_flags_aos(W.R.suffixFrontAos, A.neighIdx[(size_t)p * k + i]) >> 16) & 1u)
                validFlags |= 1u << i;
        W.validMask[p] = validFlags;
    }
    neighbor_stats_wave(W, active, count);
}

// The two shifts per valid neighbour of a queried pixel (final_gather.cu:135-177); gIdx > 0: one
// shift per neighbour onto the gather prefix's traced last vertex (final_gather.cu:248-279).
// One thread per shift of a listed pixel (list entry j, neighbour slot i, direction d), so each
// thread makes one chain of dependent record and vertex fetches instead of a pixel's k x 2 in a
// row (1.32 ms per 1080p frame with one thread per pixel).  Held to 5 waves per SIMD.
__global__ __launch_bounds__(WBLOCK) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_w_gather_setup(WaveArgs W) {
    const FrameArgs& A = W.A;
    const uint32_t k = A.k, dirs = W.gIdx == 0 ? 2u : 1u;
    const uint64_t per = (uint64_t)k * dirs;   // per > 0
    const uint64_t total = gather_total(W);
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if ((t & ~63ull) >= total * per) return;   // the whole wavefront is past the listed pixels' shifts
    const uint64_t j = t / per;
    const uint32_t p = j < total ? W.gatherList[j] : A.P;
    const uint32_t i = (uint32_t)((t % per) / dirs), d = (uint32_t)(t % dirs);
    const bool want = p < A.P && ((W.validMask[p] >> i) & 1u);
    Emit e = emit_none();
    ShState s;
    const uint32_t job = ((uint32_t)p * k + i) * 2 + d;
    if (want) {
        const uint32_t n = A.neighIdx[(size_t)p * k + i];
        const uint32_t neigh = make_ref(REF_SUFFIX_FRONT, n);
        if (W.gIdx == 0) {
            const uint32_t canon = make_ref(REF_SUFFIX_SCRATCH, p);
            e = d == 0 ? sh_start_suffix(A.S, W.R, s, neigh, canon) : sh_start_suffix(A.S, W.R, s, canon, neigh);
        } else {
            e = sh_start_suffix(A.S, W.R, s, neigh, make_ref(REF_PACKED_TRACED, p));
        }
        if (e.kind == EMIT_NONE) shift_finish(W, job, v3(0.0f), 0.0f);
    }
    const uint32_t slot = append(W.qout, e.kind);
    if (want && e.kind != EMIT_NONE) {
        store_ctx(W.qout, slot, sh_pack(s));
        emit_job(W.qout, slot, job, e);
    }
}

// FinalGatherCanon result (final_gather.cu:79-187) added to the pixel's radiance.
__global__ __launch_bounds__(WBLOCK) void k_w_gather_combine(WaveArgs W, int last) {
    const FrameArgs& A = W.A;
    const uint32_t p = gid();
    if (p >= A.P) return;
    const uint32_t pf = W.pixFlags[p];
    const uint32_t k = A.k;
    V3 rad = V3{W.radScratch[p].x, W.radScratch[p].y, W.radScratch[p].z};
    if (pf & 1u) {
        V3 contrib = v3(0.0f);
        if (W.gIdx == 0) {
            const Prefix cp = load_prefix(W.prefixScratch, A.P, p).s;
            if (!cp.valid()) {
                contrib = v3(0.0f);
            } else if (cp.nee()) {
                contrib = cp.f / cp.p;
            } else {
                const Suffix cs = load_suffix(W.suffixScratch, A.P, p).s;
                if (cs.valid()) {
                    const V3 canonContrib = cp.f * cs.f / (cp.p * cs.p);
                    if (k == 0) {
                        contrib = canonContrib;
                    } else {
                        const uint32_t validFlags = W.validMask[p];
                        const uint32_t neighCount = A.neighCount[p];
                        const float K = (float)__popc(validFlags);
                        float canonMis = 1.0f;
                        V3 suffixContrib = v3(0.0f);
                        for (uint32_t i = 0; i < neighCount; ++i) {
                            if (!((validFlags >> i) & 1u)) continue;
                            const SuffixRes nres = load_suffix_aos(W.R.suffixFrontAos, A.neighIdx[(size_t)p * k + i]);
                            const float4 a = W.shiftRes[((size_t)p * k + i) * 2];
                            const float4 b = W.shiftRes[((size_t)p * k + i) * 2 + 1];
                            const V3 fFromCanonOfNeigh = V3{a.x, a.y, a.z};
                            const float jNC = a.w;
                            const V3 fFromNeighOfCanon = V3{b.x, b.y, b.z};
                            const float jCN = b.w;
                            float nMis = neighbor_mis(fFromCanonOfNeigh, nres.s.f, jNC, K, 1.0f, nres.conf);
                            if (bad(nMis)) nMis = 0.0f;
                            const V3 nc = nMis * fFromCanonOfNeigh * nres.wSum * jNC;
                            if (!any_bad(nc)) suffixContrib = suffixContrib + nc;
                            canonMis += canonical_mis(cs.f, fFromNeighOfCanon, jCN, K, 1.0f, nres.conf);
                        }
                        V3 gatherContrib = suffixContrib * cp.f / cp.p;
                        gatherContrib = gatherContrib / (float)(__popc(validFlags) + 1);
                        V3 result = gatherContrib + (canonContrib * canonMis);
                        result = result / (K + 1.0f);
                        contrib = any_bad(result) ? v3(0.0f) : result;
                    }
                }
            }
        } else if (k > 0 && (pf & 4u) && (pf & 8u)) {   // FinalGatherNotCanon
            const uint32_t validFlags = W.validMask[p];
            const uint32_t neighCount = A.neighCount[p];
            const float K = (float)__popc(validFlags);
            V3 suffixContrib = v3(0.0f);
            for (uint32_t i = 0; i < neighCount; ++i) {
                if (!((validFlags >> i) & 1u)) continue;
                const SuffixRes nres = load_suffix_aos(W.R.suffixFrontAos, A.neighIdx[(size_t)p * k + i]);
                const float4 a = W.shiftRes[((size_t)p * k + i) * 2];
                const V3 f = V3{a.x, a.y, a.z};
                const float jNC = a.w;
                float nMis = neighbor_mis(f, nres.s.f, jNC, K, 1.0f, nres.conf);
                if (bad(nMis)) nMis = 0.0f;
                const V3 nc = nMis * f * nres.wSum * jNC;
                if (!any_bad(nc)) suffixContrib = suffixContrib + nc;
            }
            const float4 fg = W.fgScratch[p];
            V3 result = suffixContrib * V3{fg.x, fg.y, fg.z} / fg.w;
            result = result / (float)A.lp.gatherN;
            result = result / (K + 1.0f);
            contrib = any_bad(result) ? v3(0.0f) : result;
        }
        rad = rad + contrib;
    }
    W.radScratch[p] = float4{rad.x, rad.y, rad.z, 0.0f};
    if (last) {
        if (pf & 1u) {
            if (any_bad(rad)) rad = v3(0.0f);
            const uint64_t r = W.rngScratch[p];
            const uint4 g = A.gbuf[p];
            A.gbuf[p] = uint4{(uint32_t)r, (uint32_t)(r >> 32), g.z, g.w};
        }
        write_out(A, p, rad);
    }
}

// TracePrefixForFinalGather (path_gen.h:474-531) for gather prefix gIdx > 0.
__device__ __forceinline__ void fg_finish(const WaveArgs& W, uint32_t p, const FgState& s, bool ok, const Packed& last) {
    W.rngScratch[p] = s.rng;
    W.fgScratch[p] = float4{s.thr.x, s.thr.y, s.thr.z, s.p};
    W.packedScratch[2 * (size_t)p] = uint4{last.prim, fb(last.bu), fb(last.bv), fb(last.dir.x)};
    W.packedScratch[2 * (size_t)p + 1] = uint4{fb(last.dir.y), fb(last.dir.z), 0u, 0u};
    W.pixFlags[p] = (W.pixFlags[p] & 1u) | (ok ? 4u : 0u);
}

__global__ __launch_bounds__(WBLOCK) void k_w_fg_init(WaveArgs W) {
    const FrameArgs& A = W.A;
    const uint32_t p = gid();
    const bool px = p < A.P;
    Emit e = emit_none();
    FgState s;
    bool run = false;
    if (px) {
        run = (W.pixFlags[p] & 1u) != 0;
        W.pixFlags[p] &= 1u;
        if (run) {
            s.rng = W.rngScratch[p];
            s.idx = 0;
            s.thr = v3(1.0f);
            s.p = 1.0f;
            const float4 cd = W.camDir[p];
            s.dir = V3{cd.x, cd.y, cd.z};
            if (A.lp.prefixLen > 0) e = emit_closest(frame_cam(A).pos, s.dir);
            else fg_finish(W, p, s, true, packed_invalid());
        }
    }
    const uint32_t slot = append(W.qout, e.kind);
    if (run && e.kind != EMIT_NONE) {
        store_ctx(W.qout, slot, fg_pack(s));
        emit_job(W.qout, slot, p, e);
    }
}

__device__ __forceinline__ Emit fg_advance(const WaveArgs& W, uint32_t job, FgState& s, uint4 r, V3 o, V3 d) {
    const Inter it = inter_from_hit(W.A.S, r, d);
    if (!it.valid) {
        fg_finish(W, job, s, false, packed_invalid());
        return emit_none();
    }
    Rng rng;
    rng.s = s.rng;
    const BrdfSample bs = ggx_sample(W.A.S, it, rng);
    s.rng = rng.s;
    if (bs.pdf <= 0.0f) {
        fg_finish(W, job, s, false, it.pack());
        return emit_none();
    }
    s.thr = s.thr * bs.brdf;
    s.p *= bs.pdf;
    s.idx++;
    s.dir = bs.outDir;
    if (s.idx < (uint32_t)W.A.lp.prefixLen) return emit_closest(it.pos, bs.outDir);
    fg_finish(W, job, s, true, it.pack());
    return emit_none();
}

__global__ __launch_bounds__(WBLOCK) void k_w_fg_step(WaveArgs W) {
    CX_FOR_JOBS(W) {
    const uint32_t i = _base + threadIdx.x;
    const bool active = i < _n;
    Emit e = emit_none();
    FgState s;
    uint32_t job = 0;
    if (active) {
        const uint32_t slot = in_slot(W.qin, _qi, i);
        job = W.qin.job[slot];
        s = fg_unpack(load_ctx<FG_CHUNKS>(W.qin, slot));
        V3 o, d;
        get_ray(W.qin, slot, o, d);
        e = fg_advance(W, job, s, W.qin.res[slot], o, d);
    }
    const uint32_t slot = append(W.qout, e.kind);
    if (active && e.kind != EMIT_NONE) {
        store_ctx(W.qout, slot, fg_pack(s));
        emit_job(W.qout, slot, job, e);
    }
    }
}

// ======================================================================================
// Tail: a machine's last rounds in one launch
// ======================================================================================

template <int M>
struct Mach;
template <>
struct Mach<WK_PG_STEP> {
    using S = PgState;
    __device__ static S load(const WaveArgs& W, uint32_t slot) { return pg_load(W.qin, slot, W.A.lp); }
    __device__ static Emit advance(const WaveArgs& W, uint32_t job, S& s, uint4 r, V3 o, V3 d) {
        return pg_advance(W, job, s, r, o, d);
    }
};
template <>
struct Mach<WK_CP_STEP> {
    using S = CpState;
    __device__ static S load(const WaveArgs& W, uint32_t slot) { return cp_load(W.qin, slot); }
    __device__ static Emit advance(const WaveArgs& W, uint32_t job, S& s, uint4 r, V3 o, V3 d) {
        return cp_advance(W, job, s, r, o, d);
    }
};
template <>
struct Mach<WK_SG_STEP> {
    using S = SgState;
    __device__ static S load(const WaveArgs& W, uint32_t slot) {
        return sg_unpack(load_ctx<SG_CHUNKS>(W.qin, slot), W.qin.ray[slot], W.qin.ray[(size_t)W.qin.cap + slot]);
    }
    __device__ static Emit advance(const WaveArgs& W, uint32_t job, S& s, uint4 r, V3 o, V3 d) {
        return sg_advance(W, job, s, r, o, d);
    }
};
template <>
struct Mach<WK_SH_STEP> {
    using S = ShState;
    __device__ static S load(const WaveArgs& W, uint32_t slot) { return sh_unpack(load_ctx<SH_CHUNKS>(W.qin, slot)); }
    __device__ static Emit advance(const WaveArgs& W, uint32_t job, S& s, uint4 r, V3 o, V3 d) {
        return sh_advance(W, job, s, r, o, d);
    }
};
template <>
struct Mach<WK_FG_STEP> {
    using S = FgState;
    __device__ static S load(const WaveArgs& W, uint32_t slot) { return fg_unpack(load_ctx<FG_CHUNKS>(W.qin, slot)); }
    __device__ static Emit advance(const WaveArgs& W, uint32_t job, S& s, uint4 r, V3 o, V3 d) {
        return fg_advance(W, job, s, r, o, d);
    }
};

// Finishes every job of qin in one launch: each thread traces its job's queued ray (the
// traversal of k_q_trace, so the same answer), steps the job, and keeps tracing the job's next
// rays in place until it ends, instead of one traversal launch and one step launch per round.
// For a machine's last rounds, whose queues are short: every job then follows its own chain of
// rays, without waiting each round for the slowest ray of the whole queue, and the near-empty
// rounds cost nothing.  maxSteps is the machine's remaining round bound; a job still running
// after it raises the queue's overflow flag (bit 1).
template <int M, int WIDTH>
__global__ __launch_bounds__(WBLOCK) void k_w_tail(WaveArgs W, uint32_t maxSteps, int pass) {
    __shared__ int s_st[SHORT_STACK * WBLOCK];
    TravStack<SHORT_STACK> st;
    st.lds = s_st + threadIdx.x;
    st.stride = WBLOCK;
    const DScene& S = W.A.S;
    uint32_t nClosest = 0, nShadow = 0;
    CX_FOR_JOBS(W) {
    const uint32_t i = _base + threadIdx.x;
    if (i < _n) {
        const uint32_t slot = in_slot(W.qin, _qi, i);
        const uint32_t job = W.qin.job[slot];
        typename Mach<M>::S s = Mach<M>::load(W, slot);
        const float4 a = W.qin.ray[slot], b = W.qin.ray[(size_t)W.qin.cap + slot];
        Emit e;
        e.kind = i < _nc ? EMIT_CLOSEST : EMIT_SHADOW;
        e.o = V3{a.x, a.y, a.z};
        e.tmin = a.w;
        e.d = V3{b.x, b.y, b.z};
        e.tmax = b.w;
        for (uint32_t k = 0; k < maxSteps && e.kind != EMIT_NONE; k++) {
            Hit h;
            uint4 r;
            if (e.kind == EMIT_CLOSEST) {
                traverse_w<false, WIDTH>(S, e.o, e.d, e.tmin, e.tmax, h, st);
                const bool hit = h.prim != INVALID_PRIM;
                r = uint4{h.prim, fb(hit ? h.t : 0.0f), fb(hit ? h.u : 0.0f), fb(hit ? h.v : 0.0f)};
                nClosest++;
            } else {
                const bool occ = traverse_w<true, WIDTH>(S, e.o, e.d, e.tmin, e.tmax, h, st);
                r = uint4{occ ? 1u : 0u, 0u, 0u, 0u};
                nShadow++;
            }
            e = Mach<M>::advance(W, job, s, r, e.o, e.d);
        }
        if (e.kind != EMIT_NONE) atomicOr(W.qin.overflow, 2u);
    }
    }
    unsigned long long c = nClosest, sh = nShadow;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        c += __shfl_xor(c, off, 64);
        sh += __shfl_xor(sh, off, 64);
    }
    if ((threadIdx.x & 63) == 0 && W.A.counters != nullptr) {
        unsigned long long* cs =
            W.A.counters + (((blockIdx.x * blockDim.x + threadIdx.x) >> 6) % COUNTER_STRIPES) * NUM_COUNTERS;
        if (c) {
            atomicAdd(&cs[2 * pass], c);
            atomicAdd(&cs[STAT_TAIL_CLOSEST], c);
        }
        if (sh) {
            atomicAdd(&cs[2 * pass + 1], sh);
            atomicAdd(&cs[STAT_TAIL_SHADOW], sh);
        }
    }
}

// Clears the stripe counters of a queue (one lane per stripe).
__global__ void k_w_zero2(uint32_t* c) {
    if (threadIdx.x < QSTRIPES) {
        c[threadIdx.x * QCOUNT_STRIDE] = 0u;
        c[threadIdx.x * QCOUNT_STRIDE + 1] = 0u;
    }
}

// ==========================================================================================
// Host launchers
// ==========================================================================================

// Grid of the grid-stride kernels: enough wavefronts to fill every CU at full occupancy.
static dim3 grid_for(uint32_t n) { return dim3((n + WBLOCK - 1) / WBLOCK); }

// Blocks of a grid-stride launch over a queue: enough for its capacity, at most q.maxBlocks
// (the renderer's queue-grid setting, 8192 by default: 32 blocks of 128 threads per CU).
static dim3 stride_grid(const Queue& q) { return dim3(queue_grid_blocks(q.jobs, q.maxBlocks)); }

// Blocks of WBLOCK threads of `kernel` resident on the current device at once (its occupancy
// times the CU count), cached per kernel and device: the grid of the kernels that share their
// input out among their wavefronts (the traversal's chunks), whose shares shrink as the grid
// grows past it.
template <typename K>
static uint32_t resident_blocks(K kernel) {
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, uint32_t> cache;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1u << 20;
    const auto key = std::make_pair(reinterpret_cast<const void*>(kernel), dev);
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int perCu = 0, cus = 0;
    uint32_t v = 1u << 20;   // unknown: leave the queue's own grid
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kernel, WBLOCK, 0) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && perCu > 0 && cus > 0)
        v = (uint32_t)perCu * (uint32_t)cus;
    cache.emplace(key, v);
    return v;
}

template <typename K>
static dim3 range_grid(const Queue& q, K kernel) {
    return dim3(std::max(1u, std::min(stride_grid(q).x, resident_blocks(kernel))));
}

hipError_t wave_zero_counts(uint32_t* c, hipStream_t s) {
    hipLaunchKernelGGL(k_w_zero2, dim3(1), dim3(64), 0, s, c);
    return hipGetLastError();
}

hipError_t wave_trace(const DScene& S, const Queue& qin, const Queue& qout, unsigned long long* counters,
                      uint32_t minChunk, hipStream_t s) {
    const dim3 b(WBLOCK);
    const uint32_t mc = std::max(1u, std::min(minChunk, TRACE_CHUNK));
    if (S.width == 4)
        hipLaunchKernelGGL(k_q_trace<4>, range_grid(qin, k_q_trace<4>), b, 0, s, S, qin, qout, counters, mc);
    else if (S.width == 8)
        hipLaunchKernelGGL(k_q_trace<8>, range_grid(qin, k_q_trace<8>), b, 0, s, S, qin, qout, counters, mc);
    else hipLaunchKernelGGL(k_q_trace<2>, range_grid(qin, k_q_trace<2>), b, 0, s, S, qin, qout, counters, mc);
    return hipGetLastError();
}

template <int M>
static hipError_t tail_launch_w(const WaveArgs& W, uint32_t maxSteps, int pass, hipStream_t s) {
    const dim3 g = stride_grid(W.qin), b(WBLOCK);
    if (W.A.S.width == 4) hipLaunchKernelGGL((k_w_tail<M, 4>), g, b, 0, s, W, maxSteps, pass);
    else if (W.A.S.width == 8) hipLaunchKernelGGL((k_w_tail<M, 8>), g, b, 0, s, W, maxSteps, pass);
    else hipLaunchKernelGGL((k_w_tail<M, 2>), g, b, 0, s, W, maxSteps, pass);
    return hipGetLastError();
}

hipError_t wave_tail(int step, const WaveArgs& W, uint32_t maxSteps, int pass, hipStream_t s) {
    switch (step) {
        case WK_PG_STEP: return tail_launch_w<WK_PG_STEP>(W, maxSteps, pass, s);
        case WK_CP_STEP: return tail_launch_w<WK_CP_STEP>(W, maxSteps, pass, s);
        case WK_SG_STEP: return tail_launch_w<WK_SG_STEP>(W, maxSteps, pass, s);
        case WK_SH_STEP: return tail_launch_w<WK_SH_STEP>(W, maxSteps, pass, s);
        case WK_FG_STEP: return tail_launch_w<WK_FG_STEP>(W, maxSteps, pass, s);
        default: return hipErrorInvalidValue;
    }
}

// LDS of one k_w_gather_query block: GATHER_QUERY_SHORT stack entries (two ints each) per thread.
static size_t gather_query_lds(const WaveArgs&) { return (size_t)2 * GATHER_QUERY_SHORT * WBLOCK * sizeof(int); }

hipError_t wave_launch(int kernel, const WaveArgs& W, uint32_t threads, hipStream_t s) {
    const dim3 b(WBLOCK);
    switch (kernel) {
        case WK_PG_STEP: hipLaunchKernelGGL(k_w_pg_step, stride_grid(W.qin), b, 0, s, W); return hipGetLastError();
        case WK_CP_STEP: hipLaunchKernelGGL(k_w_cp_step, stride_grid(W.qin), b, 0, s, W); return hipGetLastError();
        case WK_SG_STEP: hipLaunchKernelGGL(k_w_sg_step, stride_grid(W.qin), b, 0, s, W); return hipGetLastError();
        case WK_SH_STEP: hipLaunchKernelGGL(k_w_sh_step, stride_grid(W.qin), b, 0, s, W); return hipGetLastError();
        case WK_FG_STEP: hipLaunchKernelGGL(k_w_fg_step, stride_grid(W.qin), b, 0, s, W); return hipGetLastError();
        default: break;
    }
    if (threads == 0) return hipSuccess;
    const dim3 g = grid_for(threads);
    switch (kernel) {
        case WK_PREFIX_INIT: hipLaunchKernelGGL(k_w_prefix_init, g, b, 0, s, W); break;
        case WK_PREFIX_TEMPORAL_SETUP: hipLaunchKernelGGL(k_w_prefix_temporal_setup, g, b, 0, s, W); break;
        case WK_PREFIX_RESAMPLE: hipLaunchKernelGGL(k_w_prefix_resample, g, b, 0, s, W); break;
        case WK_PREFIX_SPATIAL_SETUP: hipLaunchKernelGGL(k_w_spatial_setup<false>, g, b, 0, s, W); break;
        case WK_PREFIX_SPATIAL_RESAMPLE: hipLaunchKernelGGL(k_w_prefix_spatial_resample, g, b, 0, s, W); break;
        case WK_SUFFIX_INIT: hipLaunchKernelGGL(k_w_suffix_init, g, b, 0, s, W, 0); break;
        case WK_GATHER_SUFFIX_INIT: hipLaunchKernelGGL(k_w_suffix_init, g, b, 0, s, W, 1); break;
        case WK_SUFFIX_TEMPORAL_SETUP: hipLaunchKernelGGL(k_w_suffix_temporal_setup, g, b, 0, s, W); break;
        case WK_SUFFIX_RESAMPLE: hipLaunchKernelGGL(k_w_suffix_resample, g, b, 0, s, W); break;
        case WK_SUFFIX_SPATIAL_SETUP: hipLaunchKernelGGL(k_w_spatial_setup<true>, g, b, 0, s, W); break;
        case WK_SUFFIX_SPATIAL_RESAMPLE: hipLaunchKernelGGL(k_w_suffix_spatial_resample, g, b, 0, s, W); break;
        case WK_GATHER_INIT: hipLaunchKernelGGL(k_w_gather_init, g, b, 0, s, W); break;
        case WK_GATHER_SETUP: {
            hipLaunchKernelGGL(k_w_gather_key, g, b, 0, s, W);
            const hipError_t e = launch_key_sort(W.gatherKeys, W.gatherList, W.gatherKeysAlt, W.gatherListAlt, threads,
                                                 W.gatherCount, W.gatherSortTemp, W.gatherSortBytes, s);
            if (e != hipSuccess) return e;
            if (W.A.k <= 8) hipLaunchKernelGGL(k_w_gather_query<8>, g, b, gather_query_lds(W), s, W);
            else hipLaunchKernelGGL(k_w_gather_query<32>, g, b, gather_query_lds(W), s, W);
            const uint64_t shifts = (uint64_t)threads * W.A.k * (W.gIdx == 0 ? 2u : 1u);
            if (shifts > 0)
                hipLaunchKernelGGL(k_w_gather_setup, dim3((uint32_t)((shifts + WBLOCK - 1) / WBLOCK)), b, 0, s, W);
            break;
        }
        case WK_SUFFIX_AOS: hipLaunchKernelGGL(k_w_suffix_aos, g, b, 0, s, W); break;
        case WK_GATHER_COMBINE: hipLaunchKernelGGL(k_w_gather_combine, g, b, 0, s, W, 0); break;
        case WK_GATHER_COMBINE_LAST: hipLaunchKernelGGL(k_w_gather_combine, g, b, 0, s, W, 1); break;
        case WK_FG_INIT: hipLaunchKernelGGL(k_w_fg_init, g, b, 0, s, W); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace cx
