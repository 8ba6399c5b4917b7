// Per-pixel records and the ReSTIR path logic on the device: PrefixPath / SuffixPath /
// Reservoir / RestirGBuffer records in chunked structure-of-arrays form, path generation
// (include/graph/restir/path_gen.h), the hybrid shifts with reconnection Jacobians
// (prefix_reuse.h, suffix_reuse.h), pairwise MIS and weighted reservoir updates.
#pragma once

#include "../../include/crisoptix.h"
#include "cx_scene_dev.h"

namespace cx {

// ---- records ----

// PrefixPath + reservoir fields (PrefixPath.h:7-132, Reservoir.h:8-51).  The PCG32 of the
// record is stored as its state only (the stream is always the default one).
struct Prefix {
    uint64_t rng;
    uint32_t flags;
    float p;
    V3 f;
    float pathLen;
    V3 postReconF;
    V3 reconOutDir;
    Packed primary, recon, last;
    CX_HD uint32_t length() const { return flags & 0xFFu; }
    CX_HD void set_length(uint32_t l) { flags = (flags & ~0xFFu) | (l & 0xFFu); }
    CX_HD uint32_t recon_idx() const { return (flags >> 8) & 0xFFu; }
    CX_HD void set_recon_idx(uint32_t r) { flags = (flags & ~0xFF00u) | ((r & 0xFFu) << 8); }
    CX_HD bool valid() const { return (flags >> 16) & 1u; }
    CX_HD void set_valid(bool v) { flags = (flags & ~(1u << 16)) | (v ? (1u << 16) : 0u); }
    CX_HD bool nee() const { return (flags >> 17) & 1u; }
    CX_HD void set_nee(bool v) { flags = (flags & ~(1u << 17)) | (v ? (1u << 17) : 0u); }
};

struct PrefixRes {
    Prefix s;
    float wSum, conf;
};

struct Suffix {   // SuffixPath.h:9-104
    uint64_t rng;
    uint32_t flags;
    float p;
    V3 f;
    V3 postReconF;
    V3 reconOutDir;
    Packed recon, lastPrefix;
    CX_HD uint32_t length() const { return flags & 0xFFu; }
    CX_HD void set_length(uint32_t l) { flags = (flags & ~0xFFu) | (l & 0xFFu); }
    CX_HD uint32_t recon_idx() const { return (flags >> 8) & 0xFFu; }
    CX_HD void set_recon_idx(uint32_t r) { flags = (flags & ~0xFF00u) | ((r & 0xFFu) << 8); }
    CX_HD bool valid() const { return (flags >> 16) & 1u; }
    CX_HD void set_valid(bool v) { flags = (flags & ~(1u << 16)) | (v ? (1u << 16) : 0u); }
};

struct SuffixRes {
    Suffix s;
    float wSum, conf;
};

// Chunked SoA: chunk c of pixel p is the uint4 at base[c * P + p].
//  prefix: c0 rngLo rngHi flags p | c1 f.xyz pathLen | c2 postReconF.xyz wSum |
//          c3 reconOutDir.xyz conf | c4 primary prim bu bv dir.x | c5 primary dir.y dir.z,
//          recon prim bu | c6 recon bv dir.xyz | c7 last prim bu bv dir.x | c8 last dir.y dir.z - -
//  suffix: c0 rngLo rngHi flags p | c1 f.xyz wSum | c2 postReconF.xyz conf |
//          c3 reconOutDir.xyz recon prim | c4 recon bu bv dir.x dir.y | c5 recon dir.z,
//          last prim bu bv | c6 last dir.xyz -
//  gbuffer: c0 rngLo rngHi primaryIntValid canonFlags
static constexpr int PREFIX_CHUNKS = 9;
static constexpr int SUFFIX_CHUNKS = 7;

// canonFlags bits: the validity the later passes read from the canonical prefix / suffix
// buffers (prefix_spatial_reuse.cu:217-218, suffix_spatial_reuse.cu:171-172, final_gather.cu:293-294).
static constexpr uint32_t CANON_PREFIX_VALID = 1u;
static constexpr uint32_t CANON_PRIMARY_VALID = 2u;
static constexpr uint32_t CANON_SUFFIX_VALID = 4u;

CX_HD uint32_t fb(float f) { return __builtin_bit_cast(uint32_t, f); }
CX_HD float bf(uint32_t u) { return __builtin_bit_cast(float, u); }

__device__ __forceinline__ uint4 ld4(const uint4* base, uint32_t P, int c, uint32_t p) {
    return base[(size_t)c * P + p];
}
__device__ __forceinline__ void st4(uint4* base, uint32_t P, int c, uint32_t p, uint4 v) {
    base[(size_t)c * P + p] = v;
}

__device__ __forceinline__ uint32_t load_prefix_flags(const uint4* b, uint32_t P, uint32_t p) {
    return b[p].z;
}

__device__ __forceinline__ PrefixRes load_prefix(const uint4* b, uint32_t P, uint32_t p) {
    const uint4 c0 = ld4(b, P, 0, p), c1 = ld4(b, P, 1, p), c2 = ld4(b, P, 2, p), c3 = ld4(b, P, 3, p);
    const uint4 c4 = ld4(b, P, 4, p), c5 = ld4(b, P, 5, p), c6 = ld4(b, P, 6, p), c7 = ld4(b, P, 7, p);
    const uint4 c8 = ld4(b, P, 8, p);
    PrefixRes r;
    r.s.rng = (uint64_t)c0.x | ((uint64_t)c0.y << 32);
    r.s.flags = c0.z;
    r.s.p = bf(c0.w);
    r.s.f = V3{bf(c1.x), bf(c1.y), bf(c1.z)};
    r.s.pathLen = bf(c1.w);
    r.s.postReconF = V3{bf(c2.x), bf(c2.y), bf(c2.z)};
    r.wSum = bf(c2.w);
    r.s.reconOutDir = V3{bf(c3.x), bf(c3.y), bf(c3.z)};
    r.conf = bf(c3.w);
    r.s.primary = Packed{c4.x, bf(c4.y), bf(c4.z), V3{bf(c4.w), bf(c5.x), bf(c5.y)}};
    r.s.recon = Packed{c5.z, bf(c5.w), bf(c6.x), V3{bf(c6.y), bf(c6.z), bf(c6.w)}};
    r.s.last = Packed{c7.x, bf(c7.y), bf(c7.z), V3{bf(c7.w), bf(c8.x), bf(c8.y)}};
    return r;
}

__device__ __forceinline__ void store_prefix(uint4* b, uint32_t P, uint32_t p, const PrefixRes& r) {
    st4(b, P, 0, p, uint4{(uint32_t)r.s.rng, (uint32_t)(r.s.rng >> 32), r.s.flags, fb(r.s.p)});
    st4(b, P, 1, p, uint4{fb(r.s.f.x), fb(r.s.f.y), fb(r.s.f.z), fb(r.s.pathLen)});
    st4(b, P, 2, p, uint4{fb(r.s.postReconF.x), fb(r.s.postReconF.y), fb(r.s.postReconF.z), fb(r.wSum)});
    st4(b, P, 3, p, uint4{fb(r.s.reconOutDir.x), fb(r.s.reconOutDir.y), fb(r.s.reconOutDir.z), fb(r.conf)});
    st4(b, P, 4, p, uint4{r.s.primary.prim, fb(r.s.primary.bu), fb(r.s.primary.bv), fb(r.s.primary.dir.x)});
    st4(b, P, 5, p, uint4{fb(r.s.primary.dir.y), fb(r.s.primary.dir.z), r.s.recon.prim, fb(r.s.recon.bu)});
    st4(b, P, 6, p, uint4{fb(r.s.recon.bv), fb(r.s.recon.dir.x), fb(r.s.recon.dir.y), fb(r.s.recon.dir.z)});
    st4(b, P, 7, p, uint4{r.s.last.prim, fb(r.s.last.bu), fb(r.s.last.bv), fb(r.s.last.dir.x)});
    st4(b, P, 8, p, uint4{fb(r.s.last.dir.y), fb(r.s.last.dir.z), 0u, 0u});
}

__device__ __forceinline__ uint32_t load_suffix_flags(const uint4* b, uint32_t P, uint32_t p) { return b[p].z; }

__device__ __forceinline__ SuffixRes suffix_from_chunks(uint4 c0, uint4 c1, uint4 c2, uint4 c3, uint4 c4, uint4 c5,
                                                        uint4 c6) {
    SuffixRes r;
    r.s.rng = (uint64_t)c0.x | ((uint64_t)c0.y << 32);
    r.s.flags = c0.z;
    r.s.p = bf(c0.w);
    r.s.f = V3{bf(c1.x), bf(c1.y), bf(c1.z)};
    r.wSum = bf(c1.w);
    r.s.postReconF = V3{bf(c2.x), bf(c2.y), bf(c2.z)};
    r.conf = bf(c2.w);
    r.s.reconOutDir = V3{bf(c3.x), bf(c3.y), bf(c3.z)};
    r.s.recon = Packed{c3.w, bf(c4.x), bf(c4.y), V3{bf(c4.z), bf(c4.w), bf(c5.x)}};
    r.s.lastPrefix = Packed{c5.y, bf(c5.z), bf(c5.w), V3{bf(c6.x), bf(c6.y), bf(c6.z)}};
    return r;
}

__device__ __forceinline__ SuffixRes load_suffix(const uint4* b, uint32_t P, uint32_t p) {
    return suffix_from_chunks(ld4(b, P, 0, p), ld4(b, P, 1, p), ld4(b, P, 2, p), ld4(b, P, 3, p), ld4(b, P, 4, p),
                              ld4(b, P, 5, p), ld4(b, P, 6, p));
}

// Record-major copy of suffix reservoirs for random access: the SUFFIX_CHUNKS chunks of pixel p
// at b[p * SUFFIX_AOS_STRIDE + c], one 128-byte line per record (the last word is padding), so a
// gather that reads a neighbour's record touches one line instead of SUFFIX_CHUNKS.
static constexpr uint32_t SUFFIX_AOS_STRIDE = 8;
static_assert(SUFFIX_CHUNKS < (int)SUFFIX_AOS_STRIDE, "a suffix record fits one 128-byte line");

__device__ __forceinline__ SuffixRes load_suffix_aos(const uint4* b, uint32_t p) {
    const uint4* r = b + (size_t)p * SUFFIX_AOS_STRIDE;
    return suffix_from_chunks(r[0], r[1], r[2], r[3], r[4], r[5], r[6]);
}

__device__ __forceinline__ uint32_t load_suffix_flags_aos(const uint4* b, uint32_t p) {
    return b[(size_t)p * SUFFIX_AOS_STRIDE].z;
}

__device__ __forceinline__ void store_suffix(uint4* b, uint32_t P, uint32_t p, const SuffixRes& r) {
    st4(b, P, 0, p, uint4{(uint32_t)r.s.rng, (uint32_t)(r.s.rng >> 32), r.s.flags, fb(r.s.p)});
    st4(b, P, 1, p, uint4{fb(r.s.f.x), fb(r.s.f.y), fb(r.s.f.z), fb(r.wSum)});
    st4(b, P, 2, p, uint4{fb(r.s.postReconF.x), fb(r.s.postReconF.y), fb(r.s.postReconF.z), fb(r.conf)});
    st4(b, P, 3, p, uint4{fb(r.s.reconOutDir.x), fb(r.s.reconOutDir.y), fb(r.s.reconOutDir.z), r.s.recon.prim});
    st4(b, P, 4, p, uint4{fb(r.s.recon.bu), fb(r.s.recon.bv), fb(r.s.recon.dir.x), fb(r.s.recon.dir.y)});
    st4(b, P, 5, p, uint4{fb(r.s.recon.dir.z), r.s.lastPrefix.prim, fb(r.s.lastPrefix.bu), fb(r.s.lastPrefix.bv)});
    st4(b, P, 6, p, uint4{fb(r.s.lastPrefix.dir.x), fb(r.s.lastPrefix.dir.y), fb(r.s.lastPrefix.dir.z), 0u});
}

// ---- neighbour-search structures ----

// 30-bit Morton key of a point quantised to a 1024^3 grid: (q - lo) * scale clamped to [0, 1023]
// per axis (a NaN coordinate maps to 0), the bits interleaved x, y, z from the top.
__device__ __forceinline__ uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__device__ __forceinline__ uint32_t morton30(V3 q, V3 lo, V3 scale) {
    const float fx = fminf(fmaxf((q.x - lo.x) * scale.x, 0.0f), 1023.0f);
    const float fy = fminf(fmaxf((q.y - lo.y) * scale.y, 0.0f), 1023.0f);
    const float fz = fminf(fmaxf((q.z - lo.z) * scale.z, 0.0f), 1023.0f);
    return (expand_bits((uint32_t)fx) << 2) | (expand_bits((uint32_t)fy) << 1) | expand_bits((uint32_t)fz);
}

// Per-frame entry records before sorting (one slot per pixel / input entry).
struct EntryBuild {
    float4* pos;          // [n] pos.xyz, radius
    uint8_t* valid;       // [n]
    uint32_t* keys;       // [n] Morton key, 0xFFFFFFFF for invalid entries
    uint32_t* vals;       // [n] entry index
    uint32_t* keysAlt;    // sort outputs
    uint32_t* valsAlt;
    uint32_t* numValid;   // device scalar
    V3 lo, scale;         // Morton quantisation: (q - lo) * scale in [0, 1023]
};

// Linear BVH over the Morton-sorted valid entries (Karras, "Maximizing Parallelism in the
// Construction of BVHs, Octrees, and k-d Trees", HPG 2012): internal node i of the n - 1 covers a
// range of sorted entries whose keys share a prefix, split where the next key bit changes, so a
// node's entries form one aligned Morton cell (ties between equal keys split by slot).  Each
// internal node holds both children, 4 float4 = 64 bytes:
//   {lo0.xyz, rmax0} {hi0.xyz, ref0} {lo1.xyz, rmax1} {hi1.xyz, ref1}
// lo/hi are the exact bounds of the child's entry positions and rmax the largest entry radius
// among them; ref >= 0 is an internal node, ref < 0 a leaf range of sorted entries,
// ~ref = (first << 4) | count with count in 1..ENTRY_LEAF_MAX (a subtree of at most that many
// entries is scanned, not walked).  Node 0 is the root when more than ENTRY_LEAF_MAX entries are
// valid; with fewer the query scans them all.
static constexpr uint32_t ENTRY_LEAF_MAX = 8;
// Deepest walk: keys are 30-bit Morton codes extended by the slot index, and the common-prefix
// length of a node's range grows strictly from parent to child, so a root-to-leaf path has at most
// 62 internal nodes and a depth-first walk defers at most one sibling per node of its path.
static constexpr int ENTRY_STACK = 64;

// The bounds of a node come from a range query over an implicit complete binary tree of the
// sorted entries in groups of ENTRY_SEG_LEAF (heap order, children 2i+1 and 2i+2): each node
// holds the position bounds of its groups and the largest radius, so any range of sorted entries
// is the union of at most two groups' worth of scanned entries and two tree nodes per level.
static constexpr uint32_t ENTRY_SEG_LEAF = 8;

struct EntryTree {
    float4* entrySorted;   // [capacity] pos.xyz, radius, in key order
    uint32_t* entryPixel;  // [capacity] entry index (pixel) of each sorted slot
    float4* nodes;         // [capacity][4] internal nodes (n - 1 of them are used)
    float4* segLo;         // [2^(segLevelsCap+1) - 1] lo.xyz, largest radius of each range node
    float4* segHi;         //   hi.xyz
    uint32_t* segLevels;   // device scalar: levels of the range tree over the valid entries
    uint32_t segLevelsCap; // levels at capacity (sizes of the range-tree arrays)
    const uint32_t* sortedKeys;  // [capacity] keys in sorted order (the sort's output)
    const uint32_t* numValid;
};

// Levels of the range tree over `entries` sorted entries.
CX_HD constexpr uint32_t entry_seg_levels(uint32_t entries) {
    const uint32_t groups = entries > ENTRY_SEG_LEAF ? (entries + ENTRY_SEG_LEAF - 1) / ENTRY_SEG_LEAF : 1u;
    uint32_t l = 0;
    while ((1u << l) < groups) l++;
    return l;
}

// ---- frame parameters passed by value to every pass kernel ----

// Per-frame inputs of a frame replayed from a captured graph: the camera, the frame's random
// number and index change every frame and are read from device memory (written by a one-thread
// kernel ahead of the replay); everything else a graph bakes in stays fixed between captures.
struct FrameDyn {
    float cam[12];        // pos, U, V, W
    uint32_t random;
    uint32_t pad;
    uint64_t frameIdx;
};

struct FrameArgs {
    DScene S;
    cx_launch_params lp;
    V3 camPos, camU, camV, camW;
    uint32_t W, H, P;
    uint32_t front, back;
    uint32_t copyPrefixToBack;   // prefix_gen_temp_reuse.cu:164-169
    uint4* prefixRes[2];
    uint4* suffixRes[2];
    uint4* gbuf;
    float* out;
    EntryTree tree;
    uint32_t* neighIdx;          // [P][k]
    float* neighDist;            // [P][k]
    uint32_t* neighCount;        // [P]
    uint32_t k;
    unsigned long long* counters;   // [2 * pass + {0, 1}] closest / shadow rays, then STAT_*
    const FrameDyn* dyn;            // null: camera, random and frameIdx are the fields above
};

struct FrameCam {
    V3 pos, U, V, W;
};

__device__ __forceinline__ FrameCam frame_cam(const FrameArgs& A) {
    if (A.dyn == nullptr) return FrameCam{A.camPos, A.camU, A.camV, A.camW};
    const float* c = A.dyn->cam;
    return FrameCam{V3{c[0], c[1], c[2]}, V3{c[3], c[4], c[5]}, V3{c[6], c[7], c[8]}, V3{c[9], c[10], c[11]}};
}

__device__ __forceinline__ uint32_t frame_random(const FrameArgs& A) {
    return A.dyn == nullptr ? A.lp.random : A.dyn->random;
}

__device__ __forceinline__ uint64_t frame_index(const FrameArgs& A) {
    return A.dyn == nullptr ? A.lp.frameIdx : A.dyn->frameIdx;
}

// Counter slots after the per-pass ray counters.
static constexpr int STAT_TOTAL = 2 * CX_PASS_COUNT;
static constexpr int STAT_MIN = STAT_TOTAL + 1;
static constexpr int STAT_MAX = STAT_TOTAL + 2;
static constexpr int STAT_TAIL_CLOSEST = STAT_TOTAL + 3;   // rays traced in place by tail kernels
static constexpr int STAT_TAIL_SHADOW = STAT_TOTAL + 4;
static constexpr int NUM_COUNTERS = 32;
// The counters are replicated in 64 stripes of NUM_COUNTERS (the host sums them; min and max
// combine): device-scope atomics on one address serialise, so updates spread over stripes.
static constexpr int COUNTER_STRIPES = 64;

// Prefix-entry neighbour statistics of one pixel (PrefixAccelStruct::Stats), into the stripe
// of the calling lane.
__device__ __forceinline__ void neighbor_stat(unsigned long long* counters, uint32_t count) {
    unsigned long long* c = counters + ((blockIdx.x * blockDim.x + threadIdx.x) % COUNTER_STRIPES) * NUM_COUNTERS;
    atomicAdd(&c[STAT_TOTAL], (unsigned long long)count);
    atomicMin(&c[STAT_MIN], (unsigned long long)count);
    atomicMax(&c[STAT_MAX], (unsigned long long)count);
}

CX_HD bool is_pt(const cx_launch_params& lp) { return lp.rendererType == CX_RESTIR_PT; }
CX_HD bool is_cond(const cx_launch_params& lp) { return lp.rendererType == CX_CONDITIONAL_RESTIR; }

// ---- reservoir (Reservoir.h:8-51) ----

static constexpr float CONFIDENCE_MAX = 50.0f;

template <typename R, typename T>
__device__ __forceinline__ bool res_update(R& res, const T& sample, float w, Rng& rng, float inConf) {
    if (isnan(w) || isinf(w)) w = 0.0f;
    res.wSum += w;
    res.conf += inConf;
    res.conf = gmin(res.conf, CONFIDENCE_MAX);
    if (rng.f() < w / res.wSum) {
        res.s = sample;
        return true;
    }
    return false;
}

template <typename R>
__device__ __forceinline__ void res_finalize(R& res) {
    const float pHat = luminance(res.s.f);
    if (pHat == 0.0f) res.wSum = 0.0f;
    else res.wSum /= pHat;
}

// ---- pairwise MIS, Jacobian, neighbour selection (path_gen.h:12-72) ----

CX_HD float canonical_mis(V3 atBasis, V3 atNeighbor, float J, float K, float canonM, float neighM) {
    float w = 1.0f;
    if (luminance(atBasis) > 0.0f) {
        const float b = luminance(atBasis) * canonM;
        w = b / (b + luminance(atNeighbor) * J * neighM * K);
    }
    return w;
}

CX_HD float neighbor_mis(V3 atBasis, V3 atNeighbor, float J, float K, float canonM, float neighM) {
    float w = 0.0f;
    if (luminance(atNeighbor) > 0.0f) {
        w = luminance(atNeighbor) * neighM / (luminance(atNeighbor) * neighM + luminance(atBasis) * J * canonM / K);
    }
    return w;
}

CX_HD void select_neighbor(uint32_t px, uint32_t py, uint32_t R, Rng& rng, uint32_t& nx, uint32_t& ny) {
    const uint32_t W = 2 * R + 1;
    const uint32_t x = (px + R) - (rng.u32() % W);
    const uint32_t y = (py + R) - (rng.u32() % W);
    if (x == px && y == py) {
        nx = px + 1;
        ny = py + 1;
        return;
    }
    nx = x;
    ny = y;
}

CX_HD float recon_jacobian(V3 oldXi, V3 newXi, V3 tp, V3 tn) {
    const float term1 = fabsf(dot(normalize(newXi - tp), tn) / dot(normalize(oldXi - tp), tn));
    const float term2 = dot(tp - oldXi, tp - oldXi);
    const float term3 = dot(tp - newXi, tp - newXi);
    const float result = term1 * term2 / term3;
    if (isinf(result) || isnan(result)) return 0.0f;
    return gmax(0.0f, result);
}

// ---- path generation ----

__device__ __forceinline__ void trace_prefix(const FrameArgs& A, TraceCtx& tc, Prefix& pre, V3 origin, V3 dir,
                                             Rng& rng) {   // path_gen.h:74-250
    const DScene& S = A.S;
    const cx_launch_params& lp = A.lp;
    V3 currentPos = origin;
    V3 currentDir = dir;
    pre.flags = 0u;
    pre.pathLen = 0.0f;
    pre.rng = rng.s;
    pre.p = 1.0f;
    pre.f = v3(1.0f);
    pre.postReconF = v3(1.0f);
    pre.set_valid(true);
    Inter it;
    it.valid = false;
    it.prim = 0;
    it.bu = it.bv = 0.0f;
    it.pos = it.dir = it.nrm = v3(0.0f);
    const uint32_t maxLen = is_pt(lp) ? (uint32_t)lp.maxPathLen : (uint32_t)lp.prefixLen;
    bool postRecon = false;
    for (uint32_t traceIdx = 0; traceIdx < maxLen; ++traceIdx) {
        trace_surface(S, tc, currentPos, currentDir, 1e-3f, 1e16f, it);
        if (!it.valid) {
            pre.set_valid(false);
            return;
        }
        pre.pathLen += distance(currentPos, it.pos);
        pre.set_length(pre.length() + 1);
        if (pre.length() == 1) pre.primary = it.pack();
        if (is_cond(lp) && traceIdx == maxLen - 1) {
            if (!postRecon) {
                pre.recon = it.pack();
                pre.set_recon_idx(pre.length());
                pre.reconOutDir = v3(0.0f);
            }
            pre.set_valid(true);
            pre.last = it.pack();
            return;
        }
        const bool forceNee = is_pt(lp) && traceIdx == maxLen - 1;
        if (rng.f() < lp.neeProb || forceNee) {
            pre.p *= lp.neeProb;
            const EmitterSample es = sample_emitter(S, rng);
            const V3 lightDir = normalize(es.pos - it.pos);
            const float dist = length(es.pos - it.pos);
            if (trace_occluded(S, tc, it.pos, lightDir, 1e-3f, dist)) {
                pre.set_valid(false);
                return;
            }
            if (!postRecon) {   // path_gen.h:165-187 re-traces the same ray
                trace_surface(S, tc, currentPos, currentDir, 1e-3f, 1e16f, it);
                if (it.valid) {
                    pre.recon = it.pack();
                    pre.set_recon_idx(pre.length());
                    pre.reconOutDir = currentDir;
                    postRecon = true;
                } else {
                    pre.set_valid(false);
                    return;
                }
            }
            const BrdfEval be = ggx_eval(S, it, lightDir);
            if (be.pdf <= 0.0f) {
                pre.set_valid(false);
                return;
            }
            const float distFactor = 1.0f / (dist * dist);
            pre.f = pre.f * (be.brdf * es.color * distFactor);
            if (postRecon) pre.postReconF = pre.postReconF * (be.brdf * es.color * distFactor);
            if (pre.length() == 1) pre.f = pre.f + be.emission * distFactor;
            pre.set_nee(true);
            pre.set_valid(true);
            pre.last = it.pack();
            return;
        }
        const BrdfSample bs = ggx_sample(S, it, rng);
        if (bs.pdf <= 0.0f) {
            pre.set_valid(false);
            return;
        }
        const float reconDistance = distance(currentPos, it.pos);
        const bool canRecon =
            reconDistance > lp.reconMinDistance && bs.roughness > lp.reconMinRoughness && pre.length() >= 2;
        if (!postRecon && canRecon) {
            pre.recon = it.pack();
            pre.set_recon_idx(pre.length());
            pre.reconOutDir = bs.outDir;
            postRecon = true;
        }
        currentPos = it.pos;
        currentDir = bs.outDir;
        pre.f = pre.f * bs.brdf;
        pre.p *= bs.pdf * (1.0f - lp.neeProb);
        if (postRecon) pre.postReconF = pre.postReconF * bs.brdf;
    }
    pre.last = it.pack();
}

__device__ __forceinline__ void trace_suffix(const FrameArgs& A, TraceCtx& tc, Suffix& suf, const Prefix& pre,
                                             Rng& rng) {   // path_gen.h:252-387
    const DScene& S = A.S;
    const cx_launch_params& lp = A.lp;
    suf.f = v3(1.0f);
    suf.p = 1.0f;
    suf.postReconF = v3(1.0f);
    suf.rng = rng.s;
    suf.lastPrefix = pre.last;
    suf.flags = 0u;
    if (!pre.valid() || pre.nee()) return;
    const uint32_t maxLen = (uint32_t)lp.maxPathLen - pre.length();
    Inter it = rehydrate(S, pre.last);
    if (!it.valid) return;
    bool postRecon = false;
    V3 lastDir = v3(0.0f);
    for (uint32_t vertIdx = 0; vertIdx < maxLen; ++vertIdx) {
        const V3 lastPos = it.pos;
        if (suf.length() > 0) {
            trace_surface(S, tc, lastPos, lastDir, 1e-3f, 1e16f, it);
            if (!it.valid) return;
        }
        const bool forceNee = vertIdx == maxLen - 1;
        if (rng.f() < lp.neeProb || forceNee) {
            if (!forceNee) suf.p *= lp.neeProb;
            const EmitterSample em = sample_emitter(S, rng);
            const V3 lightVec = em.pos - it.pos;
            const float lightLen = length(lightVec);
            const V3 lightDir = normalize(lightVec);
            if (any_bad(lightDir)) return;
            if (trace_occluded(S, tc, it.pos, lightDir, 1e-3f, lightLen)) return;
            const BrdfEval be = ggx_eval(S, it, lightDir);
            if (be.pdf <= 0.0f) return;
            suf.set_length(suf.length() + 1);
            if (!postRecon) {
                Inter ri;
                ri.valid = false;
                trace_surface(S, tc, it.pos, lightDir, 1e-3f, 1e16f, ri);
                if (!ri.valid) return;
                suf.recon = ri.pack();
                suf.set_recon_idx(suf.length());
                suf.reconOutDir = v3(0.0f);
                postRecon = true;
            }
            const float distFactor = 1.0f / (lightLen * lightLen);
            suf.f = suf.f * (be.brdf * em.color * distFactor);
            suf.postReconF = suf.postReconF * (em.color * distFactor);
            if (suf.recon_idx() < suf.length()) suf.postReconF = suf.postReconF * be.brdf;
            suf.set_valid(true);
            return;
        }
        const BrdfSample bs = ggx_sample(S, it, rng);
        if (bs.pdf <= 0.0f) return;
        suf.p *= bs.pdf * (1.0f - lp.neeProb);
        suf.f = suf.f * bs.brdf;
        if (postRecon) suf.postReconF = suf.postReconF * bs.brdf;
        const float reconDistance = distance(lastPos, it.pos);
        const bool canRecon =
            reconDistance > lp.reconMinDistance && bs.roughness > lp.reconMinRoughness && pre.length() >= 1;
        if (!postRecon && canRecon) {
            suf.recon = it.pack();
            suf.set_recon_idx(pre.length());   // path_gen.h:378
            suf.reconOutDir = bs.outDir;
            postRecon = true;
        }
        lastDir = bs.outDir;
        suf.set_length(suf.length() + 1);
    }
}

__device__ __forceinline__ V3 trace_complete_path(const FrameArgs& A, TraceCtx& tc, V3 origin, V3 dir,
                                                  Rng& rng) {   // path_gen.h:389-472
    const DScene& S = A.S;
    const cx_launch_params& lp = A.lp;
    V3 currentPos = origin;
    V3 currentDir = dir;
    uint32_t depth = 0;
    V3 radiance = v3(0.0f);
    V3 thr = v3(1.0f);
    for (uint32_t traceIdx = 0; traceIdx < (uint32_t)lp.maxPathLen; ++traceIdx) {
        Inter it;
        it.valid = false;
        trace_surface(S, tc, currentPos, currentDir, 1e-3f, 1e16f, it);
        if (!it.valid) break;
        ++depth;
        if (rng.f() < lp.neeProb) {
            thr = thr / lp.neeProb;
            const EmitterSample es = sample_emitter(S, rng);
            const V3 lightDir = normalize(es.pos - it.pos);
            const float dist = length(es.pos - it.pos);
            if (trace_occluded(S, tc, it.pos, lightDir, 1e-3f, dist)) return v3(0.0f);
            const BrdfEval be = ggx_eval(S, it, lightDir);
            const float distFactor = 1.0f / (dist * dist);
            radiance = thr * be.brdf * es.color * distFactor;
            if (depth == 1) radiance = radiance + be.emission;
            break;
        }
        const BrdfSample bs = ggx_sample(S, it, rng);
        if (bs.pdf <= 0.0f) break;
        currentPos = it.pos;
        currentDir = bs.outDir;
        thr = thr * (bs.brdf / bs.pdf);
        thr = thr / (1.0f - lp.neeProb);
    }
    return radiance;
}

__device__ __forceinline__ bool trace_prefix_fg(const FrameArgs& A, TraceCtx& tc, V3& thr, float& p, Inter& last,
                                                V3 origin, V3 dir, uint32_t maxLen,
                                                Rng& rng) {   // path_gen.h:474-531
    const DScene& S = A.S;
    V3 currentPos = origin;
    V3 currentDir = dir;
    thr = v3(1.0f);
    p = 1.0f;
    last.valid = false;
    for (uint32_t i = 0; i < maxLen; ++i) {
        trace_surface(S, tc, currentPos, currentDir, 1e-3f, 1e16f, last);
        if (!last.valid) return false;
        const BrdfSample bs = ggx_sample(S, last, rng);
        if (bs.pdf <= 0.0f) return false;
        currentPos = last.pos;
        currentDir = bs.outDir;
        thr = thr * bs.brdf;
        p *= bs.pdf;
    }
    return true;
}

// ---- neighbour search over the entry tree ----

// Neighbour query under C2: every entry with |q - pos| <= r is a candidate; the k smallest
// (distance, index) are kept in ascending order.  Up to KR candidates live in registers.
template <int KR>
struct TopK {
    float d[KR];
    uint32_t i[KR];
    uint32_t count;
    __device__ __forceinline__ void init() {
#pragma unroll
        for (int j = 0; j < KR; j++) {
            d[j] = INFINITY;
            i[j] = 0xFFFFFFFFu;
        }
        count = 0;
    }
    __device__ __forceinline__ static bool less(float da, uint32_t ia, float db, uint32_t ib) {
        return da < db || (da == db && ia < ib);
    }
    // Sorted insertion; a candidate not smaller than the k-th entry changes nothing.
    __device__ __forceinline__ void insert(float cd, uint32_t ci, uint32_t k) {
#pragma unroll
        for (int j = KR - 1; j >= 0; j--) {
            if ((uint32_t)j >= k) continue;
            if (less(cd, ci, d[j], i[j])) {
                if (j > 0 && less(cd, ci, d[j - 1], i[j - 1])) {
                    d[j] = d[j - 1];
                    i[j] = i[j - 1];
                } else {
                    d[j] = cd;
                    i[j] = ci;
                }
            }
        }
        if (count < k) count++;
    }
};

// Lower bound of distance(q, p) over the positions p inside [lo, hi]: per axis the gap to the
// box (0 inside), squared and summed in the order length() uses.  Float subtraction, squaring,
// addition and sqrt are monotone, and the bounds are exact entry coordinates, so the bound never
// exceeds the float distance the leaf test computes for any entry of the node.
__device__ __forceinline__ float box_distance(V3 q, float4 lo, float4 hi) {
    const float dx = q.x < lo.x ? lo.x - q.x : (q.x > hi.x ? q.x - hi.x : 0.0f);
    const float dy = q.y < lo.y ? lo.y - q.y : (q.y > hi.y ? q.y - hi.y : 0.0f);
    const float dz = q.z < lo.z ? lo.z - q.z : (q.z > hi.z ? q.z - hi.z : 0.0f);
    return sqrtf(dx * dx + dy * dy + dz * dz);
}

// Deferred siblings of the entry query: (child reference, its distance bound).  The first SHORT
// entries live in a per-thread LDS column (two ints per entry, lds[(2 i) * stride] and
// lds[(2 i + 1) * stride]), deeper ones in a per-lane private array.
template <int SHORT>
struct QueryStack {
    int* lds;
    int stride;
    uint2 ovf[ENTRY_STACK - SHORT > 0 ? ENTRY_STACK - SHORT : 1];
    __device__ __forceinline__ void put(int i, int ref, float bound) {
        if (SHORT >= ENTRY_STACK || i < SHORT) {
            lds[(2 * i) * stride] = ref;
            lds[(2 * i + 1) * stride] = __float_as_int(bound);
        } else {
            ovf[i - SHORT] = uint2{(uint32_t)ref, __float_as_uint(bound)};
        }
    }
    __device__ __forceinline__ void get(int i, int& ref, float& bound) const {
        if (SHORT >= ENTRY_STACK || i < SHORT) {
            ref = lds[(2 * i) * stride];
            bound = __int_as_float(lds[(2 * i + 1) * stride]);
        } else {
            const uint2 v = ovf[i - SHORT];
            ref = (int)v.x;
            bound = __uint_as_float(v.y);
        }
    }
};

// Depth-first k-nearest search over the entry LBVH.  A child is entered when the distance bound
// of its positions is within the largest radius of its entries (only then can one of them have
// |q - p| <= r) and, once k candidates are held, within the k-th distance (an entry farther than
// that cannot enter the set; one at exactly that distance can, through the index tie-break).
// The nearer child (by that bound) is visited first, the other deferred with its bound and
// dropped when it is popped if the k-th distance found since has passed it.  The set found is
// the same as with an exhaustive visit of all candidates.
template <int KR, typename Stack>
__device__ __forceinline__ uint32_t entries_query(const EntryTree& T, V3 q, uint32_t k, Stack& st, uint32_t* outIdx,
                                                  float* outDist) {
    TopK<KR> top;
    top.init();
    const uint32_t n = *T.numValid;
    if (k == 0 || n == 0 || isnan(q.x) || isnan(q.y) || isnan(q.z)) return 0;
    float kth = INFINITY;   // k-th smallest distance once k candidates are held
    auto scan = [&](uint32_t first, uint32_t cnt) {
        for (uint32_t s = first; s < first + cnt; s++) {
            const float4 ep = T.entrySorted[s];
            const float dd = distance(q, V3{ep.x, ep.y, ep.z});
            if (!(dd <= ep.w) || dd > kth) continue;   // a NaN distance or radius is no candidate
            top.insert(dd, T.entryPixel[s], k);
            if (top.count == k) {
#pragma unroll
                for (int m = 0; m < KR; m++)
                    if ((uint32_t)m + 1u == k) kth = top.d[m];
            }
        }
    };
    if (n <= ENTRY_LEAF_MAX) {
        scan(0u, n);
    } else {
        int sp = 0;
        int ref = 0;
        while (true) {
            if (ref >= 0) {
                const float4* nd = T.nodes + 4 * (size_t)ref;
                const float4 lo0 = nd[0], hi0 = nd[1], lo1 = nd[2], hi1 = nd[3];
                const float b0 = box_distance(q, lo0, hi0), b1 = box_distance(q, lo1, hi1);
                const bool ok0 = b0 <= lo0.w && b0 <= kth, ok1 = b1 <= lo1.w && b1 <= kth;
                const int r0 = __float_as_int(hi0.w), r1 = __float_as_int(hi1.w);
                if (ok0 && ok1) {
                    const bool near0 = b0 <= b1;
                    st.put(sp, near0 ? r1 : r0, near0 ? b1 : b0);
                    ++sp;
                    ref = near0 ? r0 : r1;
                    continue;
                }
                if (ok0 || ok1) {
                    ref = ok0 ? r0 : r1;
                    continue;
                }
            } else {
                const uint32_t code = ~(uint32_t)ref;
                scan(code >> 4, code & 15u);
            }
            bool found = false;
            while (sp > 0 && !found) {
                --sp;
                float b;
                st.get(sp, ref, b);
                found = b <= kth;
            }
            if (!found) break;
        }
    }
#pragma unroll
    for (int j = 0; j < KR; j++) {
        if ((uint32_t)j < top.count) {
            outIdx[j] = top.i[j];
            outDist[j] = top.d[j];
        }
    }
    return top.count;
}

// ---- hybrid shifts ----

// prefix_reuse.h:10-99: shift src onto dst's primary hit; returns f in dst's domain.
__device__ __forceinline__ V3 shift_prefix(const FrameArgs& A, TraceCtx& tc, const Prefix& src, const Prefix& dst,
                                           float& J) {
    const DScene& S = A.S;
    J = 0.0f;
    if (!src.valid() || src.recon_idx() < 2) return v3(0.0f);
    const Inter srcPrimary = rehydrate(S, src.primary);
    if (!srcPrimary.valid) return v3(0.0f);
    Inter curr = rehydrate(S, dst.primary);
    if (!curr.valid) return v3(0.0f);
    Inter srcRecon = rehydrate(S, src.recon);
    if (!srcRecon.valid) return v3(0.0f);
    const int rvc = (int)(src.recon_idx() - 2u);
    const uint32_t reconVertCount = rvc > 0 ? (uint32_t)rvc : 0u;
    Rng srng;
    srng.s = src.rng;
    V3 thr = v3(1.0f);
    for (uint32_t i = 0; i < reconVertCount; ++i) {
        srng.f();
        const BrdfSample bs = ggx_sample(S, curr, srng);
        if (bs.pdf <= 0.0f) return v3(0.0f);
        thr = thr * bs.brdf;
        trace_surface(S, tc, curr.pos, bs.outDir, 1e-3f, 1e16f, curr);
        if (!curr.valid) return v3(0.0f);
    }
    const V3 reconVec = srcRecon.pos - curr.pos;
    const float reconLen = length(reconVec);
    const V3 reconDir = normalize(reconVec);
    if (isinf(reconLen) || any_bad(reconDir)) return v3(0.0f);
    if (trace_occluded(S, tc, curr.pos, reconDir, 1e-3f, reconLen)) return v3(0.0f);
    const BrdfEval b1 = ggx_eval(S, curr, reconDir);
    if (b1.pdf <= 0.0f) return v3(0.0f);
    thr = thr * b1.brdf;
    srcRecon.dir = reconDir;
    const BrdfEval b2 = ggx_eval(S, srcRecon, src.reconOutDir);
    if (b2.pdf <= 0.0f) return v3(0.0f);
    thr = thr * b2.brdf;
    J = recon_jacobian(srcPrimary.pos, curr.pos, srcRecon.pos, srcRecon.nrm);
    const V3 f = thr * src.postReconF;
    return V3{gmax(0.0f, f.x), gmax(0.0f, f.y), gmax(0.0f, f.z)};
}

// suffix_reuse.h:10-108: shift src onto the destination's last prefix vertex.
__device__ __forceinline__ V3 shift_suffix(const FrameArgs& A, TraceCtx& tc, const Suffix& src, const Inter& dstLast,
                                           float& J) {
    const DScene& S = A.S;
    J = 0.0f;
    if (!src.valid()) return v3(0.0f);
    if (!dstLast.valid) return v3(0.0f);
    const Inter srcLast = rehydrate(S, src.lastPrefix);
    if (!srcLast.valid) return v3(0.0f);
    Inter srcRecon = rehydrate(S, src.recon);
    if (!srcRecon.valid) return v3(0.0f);
    const int rvc = (int)(src.recon_idx() - 1u);
    const uint32_t reconVertCount = rvc > 0 ? (uint32_t)rvc : 0u;
    Inter curr = dstLast;
    Rng srng;
    srng.s = src.rng;
    V3 thr = v3(1.0f);
    for (uint32_t i = 0; i < reconVertCount; ++i) {
        srng.f();
        const BrdfSample bs = ggx_sample(S, curr, srng);
        if (bs.pdf <= 0.0f) return v3(0.0f);
        thr = thr * bs.brdf;
        const V3 oldPos = curr.pos;
        trace_surface(S, tc, oldPos, bs.outDir, 1e-3f, 1e16f, curr);
        if (!curr.valid) return v3(0.0f);
    }
    const V3 reconVec = srcRecon.pos - curr.pos;
    const float reconLen = length(reconVec);
    const V3 reconDir = normalize(reconVec);
    if (any_bad(reconDir)) return v3(0.0f);
    if (trace_occluded(S, tc, curr.pos, reconDir, 1e-3f, reconLen)) return v3(0.0f);
    const BrdfEval b1 = ggx_eval(S, curr, reconDir);
    if (b1.pdf <= 0.0f) return v3(0.0f);
    thr = thr * b1.brdf;
    if (src.recon_idx() < src.length()) {
        srcRecon.dir = reconDir;
        if (is_zero(src.reconOutDir)) return v3(0.0f);
        const BrdfEval b2 = ggx_eval(S, srcRecon, src.reconOutDir);
        if (b2.pdf <= 0.0f) return v3(0.0f);
        thr = thr * b2.brdf;
    }
    J = recon_jacobian(srcLast.pos, curr.pos, srcRecon.pos, srcRecon.nrm);
    return thr * src.postReconF;
}

// suffix_reuse.h:111-123
__device__ __forceinline__ V3 shift_suffix(const FrameArgs& A, TraceCtx& tc, const Suffix& src, const Suffix& dst,
                                           float& J) {
    const Inter dl = rehydrate(A.S, dst.lastPrefix);
    if (!dl.valid) return v3(0.0f);
    return shift_suffix(A, tc, src, dl, J);
}

}  // namespace cx
