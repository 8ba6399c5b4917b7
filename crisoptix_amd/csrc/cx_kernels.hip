// HIP kernels of the frame loop.  One kernel per raygen program of the reference
// (shader/*.cu), one thread per pixel, 128-thread blocks covering 16x8 pixel tiles so a
// wavefront holds a compact 16x4 screen region.  The BVH traversal stack of each thread is a
// column of a block-wide LDS array.
#include "cx_internal.h"
#include "cx_restir_dev.h"

namespace cx {

static constexpr int BLOCK = 128;
static constexpr int TILE_X = 16;
static constexpr int TILE_Y = 8;

#define CX_STACK_DECL                                  \
    __shared__ int s_stack[STACK_DEPTH * BLOCK];       \
    TraceCtx tc;                                       \
    tc.stack = s_stack + threadIdx.x;                  \
    tc.stride = BLOCK;                                 \
    tc.closest = 0;                                    \
    tc.shadow = 0;

__device__ __forceinline__ bool tile_pixel(const FrameArgs& A, uint32_t& x, uint32_t& y, uint32_t& p) {
    x = blockIdx.x * TILE_X + (threadIdx.x % TILE_X);
    y = blockIdx.y * TILE_Y + (threadIdx.x / TILE_X);
    p = y * A.W + x;
    return x < A.W && y < A.H;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One atomic per wavefront for the ray counters, into the wavefront's counter stripe;
// counters[2 * pass + {0, 1}] hold the closest-hit and shadow rays of each pass.
__device__ __forceinline__ void flush_counters(const TraceCtx& tc, unsigned long long* counters) {
    const unsigned long long c = wave_sum((unsigned long long)tc.closest);
    const unsigned long long s = wave_sum((unsigned long long)tc.shadow);
    if ((threadIdx.x & 63) == 0 && counters != nullptr) {
        unsigned long long* st =
            counters + (((blockIdx.x * blockDim.x + threadIdx.x) >> 6) % COUNTER_STRIPES) * NUM_COUNTERS;
        if (c) atomicAdd(&st[0], c);
        if (s) atomicAdd(&st[1], s);
    }
}

__device__ __forceinline__ void write_output(const FrameArgs& A, uint32_t p, V3 rad) {
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

// prefix_gen_temp_reuse.cu:122-126, Camera.h:14-22
__device__ __forceinline__ void camera_ray(const FrameArgs& A, uint32_t x, uint32_t y, Rng& rng, V3& o, V3& d) {
    const float jx = rng.f();
    const float jy = rng.f();
    float ux = ((float)x + jx) / (float)A.W;
    float uy = ((float)y + jy) / (float)A.H;
    ux = 2.0f * ux - 1.0f;
    uy = 2.0f * uy - 1.0f;
    const FrameCam c = frame_cam(A);
    o = c.pos;
    d = normalize(c.U * ux + c.V * uy + c.W);
}

__device__ __forceinline__ void store_gbuf(const FrameArgs& A, uint32_t p, uint64_t rng, uint32_t primaryValid,
                                           uint32_t canonFlags) {
    A.gbuf[p] = uint4{(uint32_t)rng, (uint32_t)(rng >> 32), primaryValid, canonFlags};
}

// ------------------------------------------------------------------------------------------
// prefix_gen_temp_reuse (shader/prefix_gen_temp_reuse.cu:11-193)
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_prefix_gen_temp(FrameArgs A) {
    CX_STACK_DECL
    uint32_t x, y, p;
    if (tile_pixel(A, x, y, p)) {
        const cx_launch_params& lp = A.lp;
        Rng rng = rng_seed(tea64(p, frame_random(A)));
        V3 origin, dir;
        camera_ray(A, x, y, rng, origin, dir);
        if (lp.rendererType == CX_CONDITIONAL_RESTIR || lp.rendererType == CX_RESTIR_PT) {
            const uint32_t oldFlags = A.gbuf[p].w;
            Prefix canon;
            canon.primary = packed_invalid();
            canon.recon = packed_invalid();
            canon.last = packed_invalid();
            canon.reconOutDir = v3(0.0f);
            trace_prefix(A, tc, canon, origin, dir, rng);
            uint32_t flags = oldFlags & CANON_SUFFIX_VALID;
            if (canon.valid()) flags |= CANON_PREFIX_VALID;
            if ((oldFlags & CANON_PRIMARY_VALID) || canon.length() >= 1) flags |= CANON_PRIMARY_VALID;
            bool validPrimary = false;
            V3 rad = v3(0.0f);
            uint4* front = A.prefixRes[A.front];
            uint4* back = A.prefixRes[A.back];
            if (canon.valid()) {   // PrefixGenTempReuse
                const float canonPHat = luminance(canon.f);
                PrefixRes curr = load_prefix(front, A.P, p);
                curr.wSum = 0.0f;
                curr.conf = 0.0f;
                bool temporal = lp.prefixEnableTemporal != 0;
                PrefixRes prev;
                if (temporal) {
                    prev = load_prefix(back, A.P, p);
                    const bool skipNee = is_cond(lp) && prev.s.nee();
                    if (prev.wSum <= 0.0f || !prev.s.valid() || skipNee) temporal = false;
                }
                if (!temporal) {
                    res_update(curr, canon, canonPHat / canon.p, rng, 1.0f);
                    res_finalize(curr);
                } else {
                    float jCanonToPrev = 0.0f, jPrevToCanon = 0.0f;
                    const V3 fFromCanonOfPrev = shift_prefix(A, tc, prev.s, canon, jPrevToCanon);
                    const V3 fFromPrevOfCanon = shift_prefix(A, tc, canon, prev.s, jCanonToPrev);
                    const float pFromCanonOfPrev = luminance(fFromCanonOfPrev) * jPrevToCanon;
                    const float K = 1.0f;
                    const float prevMis = neighbor_mis(fFromCanonOfPrev, prev.s.f, jPrevToCanon, K, 1.0f, prev.conf);
                    const float canonMis =
                        1.0f + canonical_mis(canon.f, fFromPrevOfCanon, jCanonToPrev, K, 1.0f, prev.conf);
                    const float canonRis = canonMis * canonPHat / canon.p;
                    res_update(curr, canon, canonRis, rng, 1.0f);
                    const float prevRis = prevMis * pFromCanonOfPrev * prev.wSum;
                    Prefix shifted = prev.s;   // PrefixPath.h:60-81
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
            if (is_pt(lp) && !lp.prefixEnableSpatial) write_output(A, p, validPrimary ? rad : v3(0.0f));
            store_gbuf(A, p, rng.s, validPrimary ? 1u : 0u, flags);
        } else if (lp.rendererType == CX_PATH_TRACER) {
            V3 rad = trace_complete_path(A, tc, origin, dir, rng);
            if (any_bad(rad)) rad = v3(0.0f);
            write_output(A, p, rad);
        }
    }
    flush_counters(tc, A.counters + 2 * CX_PASS_PREFIX_GEN_TEMP);
}

// ------------------------------------------------------------------------------------------
// prefix_spatial_reuse (shader/prefix_spatial_reuse.cu:11-177).  Neighbours are read from the
// back buffer, which nothing writes during the pass (pass-entry snapshot); the front->back
// copy of :162-163 is a separate copy after the kernel.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_prefix_spatial(FrameArgs A) {
    CX_STACK_DECL
    uint32_t x, y, p;
    if (tile_pixel(A, x, y, p)) {
        const cx_launch_params& lp = A.lp;
        uint4* front = A.prefixRes[A.front];
        const uint4* back = A.prefixRes[A.back];
        V3 rad = v3(0.0f);
        bool ok = false;
        const uint4 g = A.gbuf[p];
        if ((load_prefix_flags(front, A.P, p) >> 16) & 1u) {
            PrefixRes curr = load_prefix(front, A.P, p);
            const Prefix currPrefix = curr.s;
            if ((g.w & CANON_PREFIX_VALID) && (g.w & CANON_PRIMARY_VALID)) {
                Rng rng;
                rng.s = (uint64_t)g.x | ((uint64_t)g.y << 32);
                const float currWSum = curr.wSum;
                const float currConf = curr.conf;
                curr.wSum = 0.0f;
                curr.conf = 0.0f;
                const int neighCount = lp.prefixSpatialCount;
                float canonMis = 1.0f;
                uint32_t validCount = 0, validFlags = 0;
                const Rng selStart = rng;
                for (int i = 0; i < neighCount; ++i) {
                    uint32_t nx, ny;
                    select_neighbor(x, y, lp.windowRadius, rng, nx, ny);
                    if (!(nx < A.W && ny < A.H)) continue;
                    const uint32_t np = ny * A.W + nx;
                    if (!A.gbuf[np].z) continue;
                    const uint32_t nf = load_prefix_flags(back, A.P, np);
                    const bool nvalid = (nf >> 16) & 1u, nnee = (nf >> 17) & 1u;
                    if (!nvalid || (is_cond(lp) && nnee)) continue;
                    ++validCount;
                    validFlags |= 1u << i;
                }
                const float K = (float)validCount;
                Rng sel = selStart;
                for (int i = 0; i < neighCount; ++i) {
                    uint32_t nx, ny;
                    select_neighbor(x, y, lp.windowRadius, sel, nx, ny);   // replays the first loop's draws
                    if (!((validFlags >> i) & 1u)) continue;
                    const uint32_t np = ny * A.W + nx;
                    const PrefixRes nres = load_prefix(back, A.P, np);
                    float jNC = 0.0f, jCN = 0.0f;
                    const V3 fFromCanonOfNeigh = shift_prefix(A, tc, nres.s, currPrefix, jNC);
                    const float pFromCanonOfNeigh = luminance(fFromCanonOfNeigh) * jNC;
                    const V3 fFromNeighOfCanon = shift_prefix(A, tc, currPrefix, nres.s, jCN);
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
                A.gbuf[p] = uint4{(uint32_t)rng.s, (uint32_t)(rng.s >> 32), g.z, g.w};
                ok = true;
                rad = curr.wSum * curr.s.f;
            }
        }
        if (is_pt(lp) && lp.spatialRoundIdx == lp.prefixSpatialRounds - 1) write_output(A, p, ok ? rad : v3(0.0f));
    }
    flush_counters(tc, A.counters + 2 * CX_PASS_PREFIX_SPATIAL);
}

__device__ __forceinline__ void invalidate(SuffixRes& r) {   // suffix_gen_temp_reuse.cu:12-17
    r.wSum = 0.0f;
    r.conf = 0.0f;
    r.s.set_valid(false);
}

// ------------------------------------------------------------------------------------------
// suffix_gen_temp_reuse (shader/suffix_gen_temp_reuse.cu:19-134), with the unconditional
// front->back copy of :132-133 fused into the store.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_suffix_gen_temp(FrameArgs A) {
    CX_STACK_DECL
    uint32_t x, y, p;
    if (tile_pixel(A, x, y, p)) {
        const cx_launch_params& lp = A.lp;
        const uint4 g = A.gbuf[p];
        uint32_t flags = g.w;
        Rng rng;
        rng.s = (uint64_t)g.x | ((uint64_t)g.y << 32);
        uint4* front = A.suffixRes[A.front];
        uint4* back = A.suffixRes[A.back];
        SuffixRes curr = load_suffix(front, A.P, p);
        const uint32_t pf = load_prefix_flags(A.prefixRes[A.front], A.P, p);
        if (!((pf >> 16) & 1u) || ((pf >> 17) & 1u)) {
            invalidate(curr);
        } else {
            const PrefixRes pre = load_prefix(A.prefixRes[A.front], A.P, p);
            Suffix canon;
            canon.recon = packed_invalid();
            canon.reconOutDir = v3(0.0f);
            trace_suffix(A, tc, canon, pre.s, rng);
            flags = (flags & ~CANON_SUFFIX_VALID) | (canon.valid() ? CANON_SUFFIX_VALID : 0u);
            if (!canon.valid()) {
                invalidate(curr);
            } else {
                const float canonPHat = luminance(canon.f);
                curr.wSum = 0.0f;
                curr.conf = 0.0f;
                bool temporal = lp.suffixEnableTemporal != 0;
                SuffixRes prev;
                if (temporal) {
                    prev = load_suffix(back, A.P, p);
                    if (prev.wSum <= 0.0f || !prev.s.valid()) temporal = false;
                }
                if (!temporal) {
                    res_update(curr, canon, canonPHat / canon.p, rng, 1.0f);
                    res_finalize(curr);
                } else {
                    float jCanonToPrev = 0.0f, jPrevToCanon = 0.0f;
                    const V3 fFromCanonOfPrev = shift_suffix(A, tc, prev.s, canon, jPrevToCanon);
                    const V3 fFromPrevOfCanon = shift_suffix(A, tc, canon, prev.s, jCanonToPrev);
                    const float pFromCanonOfPrev = luminance(fFromCanonOfPrev) * jPrevToCanon;
                    const float K = 1.0f;
                    const float prevMis = neighbor_mis(fFromCanonOfPrev, prev.s.f, jPrevToCanon, K, 1.0f, prev.conf);
                    const float canonMis =
                        1.0f + canonical_mis(canon.f, fFromPrevOfCanon, jCanonToPrev, K, 1.0f, prev.conf);
                    const float canonRis = canonMis * canonPHat / canon.p;
                    res_update(curr, canon, canonRis, rng, 1.0f);
                    const float prevRis = prevMis * pFromCanonOfPrev * prev.wSum;
                    Suffix shifted = prev.s;   // SuffixPath.h:51-65
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
    flush_counters(tc, A.counters + 2 * CX_PASS_SUFFIX_GEN_TEMP);
}

// ------------------------------------------------------------------------------------------
// suffix_spatial_reuse (shader/suffix_spatial_reuse.cu:16-167).  Both neighbour loops read the
// back buffer (equal to the front buffer at pass entry); the copy of :165-166 follows the kernel.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_suffix_spatial(FrameArgs A) {
    CX_STACK_DECL
    uint32_t x, y, p;
    if (tile_pixel(A, x, y, p)) {
        const cx_launch_params& lp = A.lp;
        uint4* front = A.suffixRes[A.front];
        const uint4* back = A.suffixRes[A.back];
        const uint4 g = A.gbuf[p];
        SuffixRes curr = load_suffix(front, A.P, p);
        const uint32_t pf = load_prefix_flags(A.prefixRes[A.front], A.P, p);
        bool stored = false;
        if (!((pf >> 16) & 1u) || ((pf >> 17) & 1u) || !curr.s.valid() || !(g.w & CANON_SUFFIX_VALID)) {
            invalidate(curr);
        } else {
            const Suffix currSuffix = curr.s;
            Rng rng;
            rng.s = (uint64_t)g.x | ((uint64_t)g.y << 32);
            const float currWSum = curr.wSum;
            const float currConf = curr.conf;
            curr.wSum = 0.0f;
            curr.conf = 0.0f;
            const int neighCount = lp.suffixSpatialCount;
            float canonMis = 1.0f;
            uint32_t validCount = 0, validFlags = 0;
            const Rng selStart = rng;
            for (int i = 0; i < neighCount; ++i) {
                uint32_t nx, ny;
                select_neighbor(x, y, lp.windowRadius, rng, nx, ny);
                if (!(nx < A.W && ny < A.H)) continue;
                const uint32_t np = ny * A.W + nx;
                if (!A.gbuf[np].z) continue;
                if (!((load_suffix_flags(back, A.P, np) >> 16) & 1u)) continue;
                ++validCount;
                validFlags |= 1u << i;
            }
            const float K = (float)validCount;
            Rng sel = selStart;
            const Inter currLast = rehydrate(A.S, currSuffix.lastPrefix);
            for (int i = 0; i < neighCount; ++i) {
                uint32_t nx, ny;
                select_neighbor(x, y, lp.windowRadius, sel, nx, ny);
                if (!((validFlags >> i) & 1u)) continue;
                const uint32_t np = ny * A.W + nx;
                const SuffixRes nres = load_suffix(back, A.P, np);
                float jNC = 0.0f, jCN = 0.0f;
                V3 fFromCanonOfNeigh = v3(0.0f);
                if (currLast.valid) fFromCanonOfNeigh = shift_suffix(A, tc, nres.s, currLast, jNC);
                const float pFromCanonOfNeigh = luminance(fFromCanonOfNeigh) * jNC;
                const V3 fFromNeighOfCanon = shift_suffix(A, tc, currSuffix, nres.s, jCN);
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
            A.gbuf[p] = uint4{(uint32_t)rng.s, (uint32_t)(rng.s >> 32), g.z, g.w};
            stored = true;
        }
        (void)stored;
        store_suffix(front, A.P, p, curr);
    }
    flush_counters(tc, A.counters + 2 * CX_PASS_SUFFIX_SPATIAL);
}

// ------------------------------------------------------------------------------------------
// prefix_store_entries (shader/prefix_store_entries.cu:9-53) + the Morton keys of the entry
// tree build.  Entry radius = half-width of the entry's box (prefix_search_hit.cu:33-34).
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_store_entries(FrameArgs A, EntryBuild E) {
    uint32_t x, y, p;
    if (!tile_pixel(A, x, y, p)) return;
    const cx_launch_params& lp = A.lp;
    const uint32_t pf = load_prefix_flags(A.prefixRes[A.front], A.P, p);
    const uint32_t sf = load_suffix_flags(A.suffixRes[A.front], A.P, p);
    bool valid = false;
    float4 e = float4{0.0f, 0.0f, 0.0f, 0.0f};
    if (((pf >> 16) & 1u) && ((sf >> 16) & 1u)) {
        const PrefixRes pre = load_prefix(A.prefixRes[A.front], A.P, p);
        const Inter last = rehydrate(A.S, pre.s.last);
        if (last.valid) {
            float radius = lp.gatherRadius;
            if (lp.gatherRadiusType == CX_RADIUS_PATH_LENGTH) radius *= pre.s.pathLen / 4.5f;
            const float minX = last.pos.x - radius, maxX = last.pos.x + radius;
            e = float4{last.pos.x, last.pos.y, last.pos.z, (maxX - minX) / 2.0f};
            valid = true;
        }
    }
    E.pos[p] = e;
    E.valid[p] = valid ? 1 : 0;
    E.keys[p] = valid ? morton30(V3{e.x, e.y, e.z}, E.lo, E.scale) : 0xFFFFFFFFu;
    E.vals[p] = p;
    A.neighCount[p] = 0;
}

// Standalone entry upload (cx_entries_build): same record as k_store_entries.
__global__ void k_entries_from_arrays(const float* pos, const float* r, const uint8_t* valid, uint32_t n,
                                      EntryBuild E) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool v = valid[i] != 0;
    const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
    const float minX = px - r[i], maxX = px + r[i];
    const float4 e = float4{px, py, pz, (maxX - minX) / 2.0f};
    E.pos[i] = v ? e : float4{0.0f, 0.0f, 0.0f, 0.0f};
    E.valid[i] = v ? 1 : 0;
    E.keys[i] = v ? morton30(V3{px, py, pz}, E.lo, E.scale) : 0xFFFFFFFFu;
    E.vals[i] = i;
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void k_zero_u64(unsigned long long* p, uint32_t n, unsigned long long minInit) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (i % NUM_COUNTERS == (uint32_t)STAT_MIN) ? minInit : 0ull;
}

// ---- entry sort: least-significant-digit radix sort of the valid entries' Morton keys ----
//
// Four passes of 8 bits, each three launches with no host round trip and no memset (the
// element count is read on the device): per-tile digit counts, a per-digit exclusive scan over
// the tiles, and a stable scatter.  The first pass reads every slot of the entry arrays, drops
// the invalid ones (key 0xFFFFFFFF; a Morton key has 30 bits) and takes the slot index as the
// value, so it also compacts; later passes touch only the valid entries.  The result is the
// valid entries ordered by (key, slot), a deterministic order.
static constexpr int RS_THREADS = 256;
static constexpr int RS_ITEMS = 4;
static constexpr uint32_t RS_TILE = RS_THREADS * RS_ITEMS;
static constexpr int RS_WAVES = RS_THREADS / 64;

__device__ __forceinline__ uint32_t rs_count_of(bool first, uint32_t cap, const uint32_t* numValid) {
    return first ? cap : *numValid;
}

// counts[d * stride + tile] = number of elements of the tile with digit d.
__global__ __launch_bounds__(RS_THREADS) void k_rs_count(const uint32_t* keys, uint32_t cap, const uint32_t* numValid,
                                                        int first, uint32_t shift, uint32_t* counts, uint32_t stride) {
    __shared__ uint32_t h[256];
    const uint32_t n = rs_count_of(first != 0, cap, numValid);
    const uint32_t t0 = blockIdx.x * RS_TILE;
    if (t0 >= n) return;
    h[threadIdx.x] = 0u;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RS_ITEMS; i++) {
        const uint32_t idx = t0 + (uint32_t)i * RS_THREADS + threadIdx.x;
        if (idx < n) {
            const uint32_t key = keys[idx];
            if (!first || key != 0xFFFFFFFFu) atomicAdd(&h[(key >> shift) & 255u], 1u);
        }
    }
    __syncthreads();
    counts[threadIdx.x * stride + blockIdx.x] = h[threadIdx.x];
}

// Block-wide exclusive scan of one value per thread (RS_THREADS threads); returns the total.
__device__ __forceinline__ uint32_t rs_block_scan(uint32_t v, uint32_t& total, uint32_t* s_w) {
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63u) s_w[w] = x;
    __syncthreads();
    uint32_t before = 0u;
    total = 0u;
#pragma unroll
    for (int k = 0; k < RS_WAVES; k++) {
        const uint32_t c = s_w[k];
        if ((uint32_t)k < w) before += c;
        total += c;
    }
    __syncthreads();
    return before + x - v;
}

// One block per digit d: exclusive scan of counts[d][0, tiles) in place; totals[d] = the sum.
// Each thread takes RS_SCAN_ITEMS consecutive tiles, so up to 2048 tiles (2 M elements) take
// one block-wide scan instead of a chain of them.
static constexpr int RS_SCAN_ITEMS = 8;

__global__ __launch_bounds__(RS_THREADS) void k_rs_scan(uint32_t* counts, uint32_t stride, uint32_t cap,
                                                       const uint32_t* numValid, int first, uint32_t* totals) {
    __shared__ uint32_t s_w[RS_WAVES];
    const uint32_t n = rs_count_of(first != 0, cap, numValid);
    const uint32_t tiles = (n + RS_TILE - 1) / RS_TILE;
    uint32_t* row = counts + (size_t)blockIdx.x * stride;
    uint32_t carry = 0u;
    for (uint32_t b0 = 0; b0 < tiles; b0 += RS_THREADS * RS_SCAN_ITEMS) {
        const uint32_t bt = b0 + threadIdx.x * RS_SCAN_ITEMS;
        uint32_t v[RS_SCAN_ITEMS], sum = 0u;
#pragma unroll
        for (int j = 0; j < RS_SCAN_ITEMS; j++) {
            v[j] = bt + j < tiles ? row[bt + j] : 0u;
            sum += v[j];
        }
        uint32_t total;
        uint32_t run = carry + rs_block_scan(sum, total, s_w);
#pragma unroll
        for (int j = 0; j < RS_SCAN_ITEMS; j++) {
            if (bt + j < tiles) row[bt + j] = run;
            run += v[j];
        }
        carry += total;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// Stable scatter: element i of the tile goes to base[d] + (elements of digit d in earlier
// tiles) + (elements of digit d before it in this tile).  Ranks within a wavefront come from
// ballots over the digit's 8 bits; per-(item round, wavefront, digit) counts are scanned in
// element order.
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint32_t* keysIn, const uint32_t* valsIn,
                                                          uint32_t* keysOut, uint32_t* valsOut, uint32_t cap,
                                                          const uint32_t* numValid, int first, uint32_t shift,
                                                          const uint32_t* counts, uint32_t stride,
                                                          const uint32_t* totals) {
    __shared__ uint32_t s_base[256];
    __shared__ uint32_t s_cnt[RS_ITEMS * RS_WAVES * 256];
    __shared__ uint32_t s_w[RS_WAVES];
    const uint32_t n = rs_count_of(first != 0, cap, numValid);
    const uint32_t t0 = blockIdx.x * RS_TILE;
    if (t0 >= n) return;
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t dtotal;
    const uint32_t dbase = rs_block_scan(totals[threadIdx.x], dtotal, s_w);
    s_base[threadIdx.x] = dbase + counts[threadIdx.x * stride + blockIdx.x];
    // the first pass counts the valid entries: their number is the sum of the digit totals
    if (first && blockIdx.x == 0 && threadIdx.x == 0) *const_cast<uint32_t*>(numValid) = dtotal;
    for (uint32_t i = threadIdx.x; i < (uint32_t)(RS_ITEMS * RS_WAVES * 256); i += RS_THREADS) s_cnt[i] = 0u;
    __syncthreads();
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t key[RS_ITEMS], val[RS_ITEMS], dig[RS_ITEMS], rank[RS_ITEMS];
    bool ok[RS_ITEMS];
#pragma unroll
    for (int i = 0; i < RS_ITEMS; i++) {
        const uint32_t idx = t0 + (uint32_t)i * RS_THREADS + threadIdx.x;
        key[i] = idx < n ? keysIn[idx] : 0xFFFFFFFFu;
        val[i] = idx < n ? (first ? idx : valsIn[idx]) : 0u;
        ok[i] = idx < n && (!first || key[i] != 0xFFFFFFFFu);
        dig[i] = (key[i] >> shift) & 255u;
        unsigned long long m = __ballot(ok[i]);
#pragma unroll
        for (int bit = 0; bit < 8; bit++) {
            const unsigned long long b = __ballot(ok[i] && ((dig[i] >> bit) & 1u));
            m &= ((dig[i] >> bit) & 1u) ? b : ~b;
        }
        rank[i] = (uint32_t)__popcll(m & below);
        if (ok[i] && rank[i] == 0u) s_cnt[(i * RS_WAVES + w) * 256 + dig[i]] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    {
        uint32_t run = 0u;
        for (int g = 0; g < RS_ITEMS * RS_WAVES; g++) {
            const uint32_t c = s_cnt[g * 256 + threadIdx.x];
            s_cnt[g * 256 + threadIdx.x] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < RS_ITEMS; i++) {
        if (!ok[i]) continue;
        const uint32_t pos = s_base[dig[i]] + s_cnt[(i * RS_WAVES + w) * 256 + dig[i]] + rank[i];
        keysOut[pos] = key[i];
        valsOut[pos] = val[i];
    }
}

// ---- entry LBVH (Karras 2012) over the sorted valid entries ----

// Longest common prefix of sorted keys i and j extended by their slot index (equal keys compare
// their slots), or -1 outside [0, n).
__device__ __forceinline__ int lbvh_delta(const uint32_t* keys, uint32_t n, int i, int j) {
    if (j < 0 || j >= (int)n) return -1;
    const uint32_t a = keys[i], b = keys[j];
    if (a == b) return 32 + __clz((uint32_t)(i ^ j));
    return __clz(a ^ b);
}

// Sorted slot i: its entry record and pixel.
__global__ void k_lbvh_slots(EntryTree T, const uint32_t* sortedVals, const float4* entryPos) {
    const uint32_t n = *T.numValid;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t pix = sortedVals[i];
    T.entrySorted[i] = entryPos[pix];
    T.entryPixel[i] = pix;
}

__device__ __forceinline__ void grow_entry(float4& lo, float4& hi, float4 p) {
    lo = float4{fminf(lo.x, p.x), fminf(lo.y, p.y), fminf(lo.z, p.z), fmaxf(lo.w, p.w)};
    hi = float4{fmaxf(hi.x, p.x), fmaxf(hi.y, p.y), fmaxf(hi.z, p.z), 0.0f};
}

__device__ __forceinline__ void grow_box(float4& lo, float4& hi, float4 l2, float4 h2) {
    lo = float4{fminf(lo.x, l2.x), fminf(lo.y, l2.y), fminf(lo.z, l2.z), fmaxf(lo.w, l2.w)};
    hi = float4{fmaxf(hi.x, h2.x), fmaxf(hi.y, h2.y), fmaxf(hi.z, h2.z), 0.0f};
}

// Levels of the range tree over the valid entries.
__global__ void k_seg_prep(EntryTree T) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *T.segLevels = entry_seg_levels(*T.numValid);
}

// Range-tree leaf j: sorted entries [8j, 8j + 8) that are valid (an empty leaf keeps an empty box).
__global__ void k_seg_leaves(EntryTree T, uint32_t leavesCap) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t leaves = 1u << *T.segLevels;
    if (j >= leavesCap || j >= leaves) return;
    const uint32_t n = *T.numValid;
    float4 lo = float4{INFINITY, INFINITY, INFINITY, -INFINITY}, hi = float4{-INFINITY, -INFINITY, -INFINITY, 0.0f};
    const uint32_t b = j * ENTRY_SEG_LEAF, e = min(b + ENTRY_SEG_LEAF, n);
    for (uint32_t i = b; i < e; i++) grow_entry(lo, hi, T.entrySorted[i]);
    T.segLo[leaves - 1u + j] = lo;
    T.segHi[leaves - 1u + j] = hi;
}

// Range-tree nodes of level l (heap indices [2^l - 1, 2^(l+1) - 1)) above the leaf level.
__global__ void k_seg_level(EntryTree T, uint32_t level, uint32_t count) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count || level >= *T.segLevels) return;
    const uint32_t node = (1u << level) - 1u + j;
    float4 lo = T.segLo[2 * node + 1], hi = T.segHi[2 * node + 1];
    grow_box(lo, hi, T.segLo[2 * node + 2], T.segHi[2 * node + 2]);
    T.segLo[node] = lo;
    T.segHi[node] = hi;
}

// Bounds (and largest radius) of the sorted entries [a, b], b >= a: short ranges are scanned;
// otherwise the partial groups at both ends are scanned and the whole groups between them are
// covered bottom-up by the classic iterative segment-tree walk (at most two nodes per level).
__device__ __forceinline__ void range_bounds(const EntryTree& T, uint32_t levels, uint32_t a, uint32_t b, float4& lo,
                                             float4& hi) {
    lo = float4{INFINITY, INFINITY, INFINITY, -INFINITY};
    hi = float4{-INFINITY, -INFINITY, -INFINITY, 0.0f};
    const uint32_t ga = a / ENTRY_SEG_LEAF, gb = b / ENTRY_SEG_LEAF;
    if (gb - ga <= 1u) {
        for (uint32_t i = a; i <= b; i++) grow_entry(lo, hi, T.entrySorted[i]);
        return;
    }
    for (uint32_t i = a; i < (ga + 1u) * ENTRY_SEG_LEAF; i++) grow_entry(lo, hi, T.entrySorted[i]);
    for (uint32_t i = gb * ENTRY_SEG_LEAF; i <= b; i++) grow_entry(lo, hi, T.entrySorted[i]);
    // whole groups [ga + 1, gb) as 1-based heap indices [l, r)
    uint32_t l = (1u << levels) + ga + 1u, r = (1u << levels) + gb;
    while (l < r) {
        if (l & 1u) {
            grow_box(lo, hi, T.segLo[l - 1u], T.segHi[l - 1u]);
            l++;
        }
        if (r & 1u) {
            --r;
            grow_box(lo, hi, T.segLo[r - 1u], T.segHi[r - 1u]);
        }
        l >>= 1;
        r >>= 1;
    }
}

// Internal node i: its key range and split (the paper's determineRange and findSplit), then the
// complete 64-byte record: each child's bounds and largest radius from range_bounds, and its
// reference (a range of at most ENTRY_LEAF_MAX entries becomes a leaf range).  Nodes are
// independent, so the build needs no synchronisation between threads.
__global__ void k_lbvh_nodes(EntryTree T) {
    const uint32_t n = *T.numValid;
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (n < 2 || i >= (int)n - 1) return;
    const uint32_t* K = T.sortedKeys;
    const int d = lbvh_delta(K, n, i, i + 1) > lbvh_delta(K, n, i, i - 1) ? 1 : -1;
    const int dmin = lbvh_delta(K, n, i, i - d);
    int lmax = 2;
    while (lbvh_delta(K, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (lbvh_delta(K, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = lbvh_delta(K, n, i, j);
    int s = 0;
    for (int t = (l + 1) / 2;; t = (t + 1) / 2) {
        if (lbvh_delta(K, n, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
    }
    const int g = i + s * d + (d < 0 ? -1 : 0);
    const int first = min(i, j), last = max(i, j);
    const uint32_t nl = (uint32_t)(g - first + 1), nr = (uint32_t)(last - g);
    const int refL = nl <= ENTRY_LEAF_MAX ? (int)~(((uint32_t)first << 4) | nl) : g;
    const int refR = nr <= ENTRY_LEAF_MAX ? (int)~(((uint32_t)(g + 1) << 4) | nr) : g + 1;
    const uint32_t levels = *T.segLevels;
    float4 lo0, hi0, lo1, hi1;
    range_bounds(T, levels, (uint32_t)first, (uint32_t)g, lo0, hi0);
    range_bounds(T, levels, (uint32_t)g + 1u, (uint32_t)last, lo1, hi1);
    float4* nd = T.nodes + 4 * (size_t)i;
    nd[0] = lo0;
    nd[1] = float4{hi0.x, hi0.y, hi0.z, __int_as_float(refL)};
    nd[2] = lo1;
    nd[3] = float4{hi1.x, hi1.y, hi1.z, __int_as_float(refR)};
}

template <int KR>
__global__ __launch_bounds__(BLOCK) void k_entries_query(EntryTree T, const float* q, uint32_t nq, uint32_t k,
                                                        uint32_t* idx, float* dist, uint32_t* count) {
    __shared__ int s_stack[STACK_DEPTH * BLOCK];
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nq) return;
    for (uint32_t i = 0; i < k; i++) {
        idx[(size_t)j * k + i] = 0xFFFFFFFFu;
        dist[(size_t)j * k + i] = 0.0f;
    }
    QueryStack<STACK_DEPTH / 2> st;
    st.lds = s_stack + threadIdx.x;
    st.stride = BLOCK;
    count[j] = entries_query<KR>(T, V3{q[3 * j], q[3 * j + 1], q[3 * j + 2]}, k, st, idx + (size_t)j * k,
                                 dist + (size_t)j * k);
}

// ------------------------------------------------------------------------------------------
// final_gather (shader/final_gather.cu:69-356)
// ------------------------------------------------------------------------------------------
template <int KR>
__device__ __forceinline__ V3 gather_canon(const FrameArgs& A, TraceCtx& tc, uint32_t p, V3 origin, V3 dir,
                                           uint32_t k, Rng& rng) {
    Prefix cp;
    cp.primary = packed_invalid();
    cp.recon = packed_invalid();
    cp.last = packed_invalid();
    cp.reconOutDir = v3(0.0f);
    cp.rng = PCG_DEFAULT_STATE;
    trace_prefix(A, tc, cp, origin, dir, rng);
    if (!cp.valid()) return v3(0.0f);
    if (cp.nee()) return cp.f / cp.p;
    Suffix cs;
    cs.recon = packed_invalid();
    cs.reconOutDir = v3(0.0f);
    trace_suffix(A, tc, cs, cp, rng);
    if (!cs.valid()) return v3(0.0f);
    const V3 canonContrib = cp.f * cs.f / (cp.p * cs.p);
    if (k == 0) return canonContrib;
    const Inter last = rehydrate(A.S, cs.lastPrefix);
    uint32_t* nIdx = A.neighIdx + (size_t)p * k;
    float* nDist = A.neighDist + (size_t)p * k;
    QueryStack<STACK_DEPTH / 2> qst;
    qst.lds = tc.stack;
    qst.stride = tc.stride;
    const uint32_t neighCount = entries_query<KR>(A.tree, last.pos, k, qst, nIdx, nDist);
    A.neighCount[p] = neighCount;
    if (A.lp.trackPrefixStats) neighbor_stat(A.counters, neighCount);
    const uint4* sfront = A.suffixRes[A.front];
    uint32_t validCount = 0, validFlags = 0;
    for (uint32_t i = 0; i < neighCount; ++i) {
        if (!((load_suffix_flags(sfront, A.P, nIdx[i]) >> 16) & 1u)) continue;
        ++validCount;
        validFlags |= 1u << i;
    }
    V3 suffixContrib = v3(0.0f);
    const float K = (float)validCount;
    float canonMis = 1.0f;
    for (uint32_t i = 0; i < neighCount; ++i) {
        if (!(validFlags & (1u << i))) continue;
        const SuffixRes nres = load_suffix(sfront, A.P, nIdx[i]);
        float jNC = 0.0f, jCN = 0.0f;
        V3 fFromCanonOfNeigh = v3(0.0f);
        if (last.valid) fFromCanonOfNeigh = shift_suffix(A, tc, nres.s, last, jNC);
        const V3 fFromNeighOfCanon = shift_suffix(A, tc, cs, nres.s, jCN);
        float nMis = neighbor_mis(fFromCanonOfNeigh, nres.s.f, jNC, K, 1.0f, nres.conf);
        if (bad(nMis)) nMis = 0.0f;
        const V3 nc = nMis * fFromCanonOfNeigh * nres.wSum * jNC;
        if (!any_bad(nc)) suffixContrib = suffixContrib + nc;
        canonMis += canonical_mis(cs.f, fFromNeighOfCanon, jCN, K, 1.0f, nres.conf);
    }
    V3 gatherContrib = suffixContrib * cp.f / cp.p;
    gatherContrib = gatherContrib / (float)(validCount + 1);
    V3 result = gatherContrib + (canonContrib * canonMis);
    result = result / (K + 1.0f);
    if (any_bad(result)) return v3(0.0f);
    return result;
}

template <int KR>
__device__ __forceinline__ V3 gather_not_canon(const FrameArgs& A, TraceCtx& tc, uint32_t p, V3 origin, V3 dir,
                                               uint32_t k, Rng& rng) {
    if (k == 0) return v3(0.0f);
    Inter last;
    last.valid = false;
    V3 thr = v3(0.0f);
    float pp = 0.0f;
    if (!trace_prefix_fg(A, tc, thr, pp, last, origin, dir, (uint32_t)A.lp.prefixLen, rng)) return v3(0.0f);
    uint32_t* nIdx = A.neighIdx + (size_t)p * k;
    float* nDist = A.neighDist + (size_t)p * k;
    QueryStack<STACK_DEPTH / 2> qst;
    qst.lds = tc.stack;
    qst.stride = tc.stride;
    const uint32_t neighCount = entries_query<KR>(A.tree, last.pos, k, qst, nIdx, nDist);
    A.neighCount[p] = neighCount;
    if (A.lp.trackPrefixStats) neighbor_stat(A.counters, neighCount);
    const uint4* sfront = A.suffixRes[A.front];
    uint32_t validCount = 0, validFlags = 0;
    for (uint32_t i = 0; i < neighCount; ++i) {
        if (!((load_suffix_flags(sfront, A.P, nIdx[i]) >> 16) & 1u)) continue;
        ++validCount;
        validFlags |= 1u << i;
    }
    V3 suffixContrib = v3(0.0f);
    const float K = (float)validCount;
    for (uint32_t i = 0; i < neighCount; ++i) {
        if (!(validFlags & (1u << i))) continue;
        const SuffixRes nres = load_suffix(sfront, A.P, nIdx[i]);
        float jNC = 0.0f;
        const V3 f = shift_suffix(A, tc, nres.s, last, jNC);
        float nMis = neighbor_mis(f, nres.s.f, jNC, K, 1.0f, nres.conf);
        if (bad(nMis)) nMis = 0.0f;
        const V3 nc = nMis * f * nres.wSum * jNC;
        if (!any_bad(nc)) suffixContrib = suffixContrib + nc;
    }
    V3 result = suffixContrib * thr / pp;
    result = result / (float)A.lp.gatherN;
    result = result / (K + 1.0f);
    if (any_bad(result)) return v3(0.0f);
    return result;
}

template <int KR>
__global__ __launch_bounds__(BLOCK) void k_final_gather(FrameArgs A) {
    CX_STACK_DECL
    uint32_t x, y, p;
    if (tile_pixel(A, x, y, p)) {
        const uint4 g = A.gbuf[p];
        V3 rad = v3(0.0f);
        if (g.w & CANON_PRIMARY_VALID) {
            Rng rng;
            rng.s = (uint64_t)g.x | ((uint64_t)g.y << 32);
            V3 origin, dir;
            camera_ray(A, x, y, rng, origin, dir);
            const uint32_t k = A.k;
            rad = rad + gather_canon<KR>(A, tc, p, origin, dir, k, rng);
            for (int gi = 1; gi < A.lp.gatherN; ++gi) rad = rad + gather_not_canon<KR>(A, tc, p, origin, dir, k, rng);
            if (any_bad(rad)) rad = v3(0.0f);
            A.gbuf[p] = uint4{(uint32_t)rng.s, (uint32_t)(rng.s >> 32), g.z, g.w};
        }
        write_output(A, p, rad);
    }
    flush_counters(tc, A.counters + 2 * CX_PASS_FINAL_GATHER);
}

// ------------------------------------------------------------------------------------------
// State reset (the uploads of src/Renderer.cpp:155-199)
// ------------------------------------------------------------------------------------------
__global__ void k_reset_state(FrameArgs A) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.P) return;
    const uint64_t defRng = rng_seed(PCG_DEFAULT_STATE).s;
    PrefixRes pr;
    pr.s.rng = defRng;
    pr.s.flags = 0;
    pr.s.p = 0.0f;
    pr.s.f = v3(0.0f);
    pr.s.pathLen = 0.0f;
    pr.s.postReconF = v3(0.0f);
    pr.s.reconOutDir = v3(0.0f);
    pr.s.primary = packed_invalid();
    pr.s.recon = packed_invalid();
    pr.s.last = packed_invalid();
    pr.wSum = 0.0f;
    pr.conf = 0.0f;
    SuffixRes sr;
    sr.s.rng = defRng;
    sr.s.flags = 0;
    sr.s.p = 0.0f;
    sr.s.f = v3(0.0f);
    sr.s.postReconF = v3(0.0f);
    sr.s.reconOutDir = v3(0.0f);
    sr.s.recon = packed_invalid();
    sr.s.lastPrefix = packed_invalid();
    sr.wSum = 0.0f;
    sr.conf = 0.0f;
    for (int b = 0; b < 2; b++) {
        store_prefix(A.prefixRes[b], A.P, p, pr);
        store_suffix(A.suffixRes[b], A.P, p, sr);
    }
    A.gbuf[p] = uint4{(uint32_t)defRng, (uint32_t)(defRng >> 32), 0u, 0u};
    A.neighCount[p] = 0;
}

// ------------------------------------------------------------------------------------------
// Standalone traversal (cx_trace_closest / cx_trace_occluded)
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(BLOCK) void k_trace(DScene S, const float* rays, uint32_t n, int anyHit,
                                                 uint32_t* prim, float* t, float* u, float* v, uint32_t* occ) {
    __shared__ int s_stack[STACK_DEPTH * BLOCK];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = rays + 8 * (size_t)i;
    Hit h;
    const V3 o = V3{r[0], r[1], r[2]}, d = V3{r[3], r[4], r[5]};
    if (anyHit) {
        occ[i] = traverse<true>(S, o, d, r[6], r[7], h, s_stack + threadIdx.x, BLOCK) ? 1u : 0u;
    } else {
        traverse<false>(S, o, d, r[6], r[7], h, s_stack + threadIdx.x, BLOCK);
        const bool hit = h.prim != INVALID_PRIM;
        prim[i] = h.prim;
        t[i] = hit ? h.t : 0.0f;
        u[i] = hit ? h.u : 0.0f;
        v[i] = hit ? h.v : 0.0f;
    }
}

// ==========================================================================================
// Host launchers
// ==========================================================================================

static dim3 tile_grid(uint32_t W, uint32_t H) { return dim3((W + TILE_X - 1) / TILE_X, (H + TILE_Y - 1) / TILE_Y); }

hipError_t launch_pass(int pass, const FrameArgs& A, hipStream_t s) {
    const dim3 grid = tile_grid(A.W, A.H);
    switch (pass) {
        case CX_PASS_PREFIX_GEN_TEMP: hipLaunchKernelGGL(k_prefix_gen_temp, grid, dim3(BLOCK), 0, s, A); break;
        case CX_PASS_PREFIX_SPATIAL: hipLaunchKernelGGL(k_prefix_spatial, grid, dim3(BLOCK), 0, s, A); break;
        case CX_PASS_SUFFIX_GEN_TEMP: hipLaunchKernelGGL(k_suffix_gen_temp, grid, dim3(BLOCK), 0, s, A); break;
        case CX_PASS_SUFFIX_SPATIAL: hipLaunchKernelGGL(k_suffix_spatial, grid, dim3(BLOCK), 0, s, A); break;
        case CX_PASS_FINAL_GATHER:
            if (A.k <= 8) hipLaunchKernelGGL(k_final_gather<8>, grid, dim3(BLOCK), 0, s, A);
            else hipLaunchKernelGGL(k_final_gather<32>, grid, dim3(BLOCK), 0, s, A);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_reset(const FrameArgs& A, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_state, dim3((A.P + 255) / 256), dim3(256), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_zero_counters(unsigned long long* c, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u64, dim3((n + 255) / 256), dim3(256), 0, s, c, n, 0xFFFFFFFFull);
    return hipGetLastError();
}

__global__ void k_set_dyn(FrameDyn* dst, FrameDyn d) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *dst = d;
}

// The per-frame inputs of a graph replay travel as a kernel argument, so the host's copy is
// free as soon as the launch returns.
hipError_t launch_set_dyn(FrameDyn* dst, const FrameDyn& d, hipStream_t s) {
    hipLaunchKernelGGL(k_set_dyn, dim3(1), dim3(64), 0, s, dst, d);
    return hipGetLastError();
}

// Device-random frames: the camera and frame index come as arguments, the frame's random number
// from a PCG32 whose state stays in device memory (the host generator's step, include/util/
// random.h:52-68), so a frame recorded into a graph draws a new number on every replay.
__global__ void k_next_dyn(FrameDyn* dst, uint64_t* state, FrameDyn d) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const uint64_t old = *state;
    *state = old * PCG_MULT + PCG_INC;
    const uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
    const uint32_t rot = (uint32_t)(old >> 59);
    d.random = (xs >> rot) | (xs << ((32u - rot) & 31u));
    *dst = d;
}

hipError_t launch_next_dyn(FrameDyn* dst, uint64_t* state, const FrameDyn& d, hipStream_t s) {
    hipLaunchKernelGGL(k_next_dyn, dim3(1), dim3(64), 0, s, dst, state, d);
    return hipGetLastError();
}

// Material table rows from device memory (cx_material layout, 10 floats); texture bindings stay.
__global__ void k_set_materials(DMaterial* dst, const float* rows, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = rows + 10 * (size_t)i;
    dst[i].diffRough = make_float4(r[0], r[1], r[2], r[9]);
    dst[i].spec = make_float4(r[3], r[4], r[5], 0.0f);
    dst[i].emissive = make_float4(r[6], r[7], r[8], 0.0f);
}

hipError_t launch_set_materials(DMaterial* dst, const float* rows, uint32_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_set_materials, dim3((n + 255) / 256), dim3(256), 0, s, dst, rows, n);
    return hipGetLastError();
}

hipError_t launch_store_entries(const FrameArgs& A, const EntryBuild& E, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, E.numValid, 1u);
    hipLaunchKernelGGL(k_store_entries, tile_grid(A.W, A.H), dim3(BLOCK), 0, s, A, E);
    return hipGetLastError();
}

hipError_t launch_entries_from_arrays(const float* pos, const float* r, const uint8_t* valid, uint32_t n,
                                      const EntryBuild& E, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, E.numValid, 1u);
    if (n) hipLaunchKernelGGL(k_entries_from_arrays, dim3((n + 255) / 256), dim3(256), 0, s, pos, r, valid, n, E);
    return hipGetLastError();
}

// Scratch of the entry sort: per-digit tile counts [256][tiles] and the digit totals.
size_t sort_temp_bytes(uint32_t n) {
    const size_t tiles = (std::max<uint32_t>(n, 1u) + RS_TILE - 1) / RS_TILE;
    return (256 * tiles + 256) * sizeof(uint32_t);
}

// Sorts the n slots of keys (0xFFFFFFFF = invalid, dropped) by key, stably: afterwards keys[0,
// *numValid) are the valid keys in order and vals[] their slot indices.  keysAlt / valsAlt are
// the ping-pong buffers; temp holds sort_temp_bytes(n).
hipError_t launch_key_sort(uint32_t* keys, uint32_t* vals, uint32_t* keysAlt, uint32_t* valsAlt, uint32_t n,
                           uint32_t* numValid, void* temp, size_t tempBytes, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (sort_temp_bytes(n) > tempBytes) return hipErrorInvalidValue;
    const uint32_t tiles = (n + RS_TILE - 1) / RS_TILE;
    uint32_t* counts = (uint32_t*)temp;
    uint32_t* totals = counts + (size_t)256 * tiles;
    uint32_t* kIn = keys;
    uint32_t* vIn = vals;
    uint32_t* kOut = keysAlt;
    uint32_t* vOut = valsAlt;
    for (int pass = 0; pass < 4; pass++) {
        const int first = pass == 0 ? 1 : 0;
        const uint32_t shift = 8u * (uint32_t)pass;
        hipLaunchKernelGGL(k_rs_count, dim3(tiles), dim3(RS_THREADS), 0, s, kIn, n, numValid, first, shift, counts,
                           tiles);
        hipLaunchKernelGGL(k_rs_scan, dim3(256), dim3(RS_THREADS), 0, s, counts, tiles, n, numValid, first, totals);
        hipLaunchKernelGGL(k_rs_scatter, dim3(tiles), dim3(RS_THREADS), 0, s, kIn, vIn, kOut, vOut, n, numValid,
                           first, shift, counts, tiles, totals);
        std::swap(kIn, kOut);
        std::swap(vIn, vOut);
    }
    // four passes: the sorted data is back in keys / vals
    return hipGetLastError();
}

// Sorts the valid entries by Morton key and builds the LBVH over them: the sorted slots, the
// range tree (one launch per level; levels above the valid entries' tree return at once), then
// every internal node at once.  Grids cover the capacity n; kernels read the valid count on the
// device.
hipError_t launch_tree_build(const EntryBuild& E, const EntryTree& T, uint32_t n, void* temp, size_t tempBytes,
                             hipStream_t s) {
    if (n == 0) return hipGetLastError();   // numValid was cleared when the entries were written
    const hipError_t e = launch_key_sort(E.keys, E.vals, E.keysAlt, E.valsAlt, n, E.numValid, temp, tempBytes, s);
    if (e != hipSuccess) return e;
    const dim3 g((n + 255) / 256), b(256);
    hipLaunchKernelGGL(k_lbvh_slots, g, b, 0, s, T, E.vals, E.pos);
    hipLaunchKernelGGL(k_seg_prep, dim3(1), dim3(64), 0, s, T);
    const uint32_t leavesCap = 1u << T.segLevelsCap;
    hipLaunchKernelGGL(k_seg_leaves, dim3((leavesCap + 255) / 256), b, 0, s, T, leavesCap);
    for (int l = (int)T.segLevelsCap - 1; l >= 0; l--) {
        const uint32_t cnt = 1u << l;
        hipLaunchKernelGGL(k_seg_level, dim3((cnt + 255) / 256), b, 0, s, T, (uint32_t)l, cnt);
    }
    hipLaunchKernelGGL(k_lbvh_nodes, g, b, 0, s, T);
    return hipGetLastError();
}

hipError_t launch_entries_query(const EntryTree& T, const float* q, uint32_t nq, uint32_t k, uint32_t* idx,
                                float* dist, uint32_t* count, hipStream_t s) {
    if (nq == 0) return hipSuccess;
    const dim3 grid((nq + BLOCK - 1) / BLOCK);
    if (k <= 8) hipLaunchKernelGGL(k_entries_query<8>, grid, dim3(BLOCK), 0, s, T, q, nq, k, idx, dist, count);
    else hipLaunchKernelGGL(k_entries_query<32>, grid, dim3(BLOCK), 0, s, T, q, nq, k, idx, dist, count);
    return hipGetLastError();
}

hipError_t launch_trace(const DScene& S, const float* rays, uint32_t n, int anyHit, uint32_t* prim, float* t,
                        float* u, float* v, uint32_t* occ, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_trace, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, s, S, rays, n, anyHit, prim, t, u, v,
                       occ);
    return hipGetLastError();
}

}  // namespace cx
