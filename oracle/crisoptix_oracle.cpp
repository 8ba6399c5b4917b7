// ORACLE -- test infrastructure only.
//
// A plain host C++17 restatement of the CrisOptix ReSTIR PT / Conditional ReSTIR shader
// logic, used as the checker of the HIP kernels (tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py).  It is never linked into, loaded by or called from the
// product library.  Every function cites the reference file:line it restates.
//
// Canonicalisations (shared with the HIP kernels, documented in the repository notes):
//  C1 spatial passes read neighbour reservoirs as they were at pass entry;
//  C2 neighbour set = the k smallest (distance, pixelIdx) among valid in-radius entries;
//  C3 Next2d draws x before y, NextUint64 takes the low word from the first draw;
//  C4 the per-frame random comes from a seeded PCG32;
//  C5 watertight ray/triangle test, closest hit = min (t, primitive), [tmin, tmax] inclusive;
//  C6 sin/cos/pow are the project-written functions of orc_math.h;
//  C7 rehydrated normals are not normalised; C8 one instance, identity transform;
//  C9' emitter triangles are traced geometry only where the scene lists them, and then only
//      closest-hit rays see them: occlusion rays pass over triangles without CX_VIS_OCCLUSION
//      (an OptiX visibility mask), so a shadow ray ending on a light is not occluded by it.
//
// Parity pins (tests/test_oracle_*.py): PCG32 against the PCG reference demo output and an
// independent restatement from the PCG paper, TEA against the published TEA test vector,
// GGX closed forms and Monte-Carlo normalisation, brute-force closest hit and occlusion,
// brute-force neighbour sets, closed forms of the Jacobian and the pairwise MIS weights, the
// pairwise weights' partition of unity for reciprocal Jacobians, the prefix identity shift and
// the suffix identity shift at an NEE reconnection and at a BSDF reconnection on the first
// suffix vertex with prefix length 1 (J = 1, the suffix's own value), the final gather's
// combination in its limits (unreachable neighbours give the canonical estimate), a converged
// FP64 reference of the path-tracer estimator, and the canonical prefix + suffix chain of
// Conditional ReSTIR against that FP64 reference one vertex shorter.  Shifts whose reconIdx
// (path_gen.h:378: the prefix length) differs from the reconnection vertex's index stay "parity
// unpinned" against the reference itself (it cannot be built here: no OptiX SDK, CUDA toolkit or
// glm); no identity fixes their values.

#include "orc_math.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/crisoptix.h"

namespace orc {

static constexpr float PI = 3.14159265358979323846264f;   // shader/brdf.cu:12
static constexpr float CONFIDENCE_MAX = 50.0f;            // include/graph/restir/Reservoir.h:6
static constexpr uint32_t INVALID_PRIM = 0xFFFFFFFFu;

// ---------------------------------------------------------------------------------------
// Random numbers (include/util/random.h)
// ---------------------------------------------------------------------------------------

// TEA with a caller key, 'rounds' rounds.  The reference's SampleTEA32/64
// (include/util/random.h:7-38) is this function with key {0xa341316c, 0xc8013ea4,
// 0xad90777d, 0x7e95761e}; the general form is kept so the published TEA test vector can
// pin it.
static inline void tea(uint32_t& v0, uint32_t& v1, const uint32_t key[4], uint32_t rounds) {
    uint32_t s0 = 0;
    for (uint32_t n = 0; n < rounds; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + key[0]) ^ (v1 + s0) ^ ((v1 >> 5) + key[1]);
        v1 += ((v0 << 4) + key[2]) ^ (v0 + s0) ^ ((v0 >> 5) + key[3]);
    }
}
static const uint32_t TEA_KEY[4] = {0xa341316cu, 0xc8013ea4u, 0xad90777du, 0x7e95761eu};

static inline uint32_t sample_tea32(uint32_t a, uint32_t b, uint32_t rounds = 4) {
    tea(a, b, TEA_KEY, rounds);
    return a;
}
static inline uint64_t sample_tea64(uint32_t a, uint32_t b, uint32_t rounds = 4) {
    tea(a, b, TEA_KEY, rounds);
    return (uint64_t)a | ((uint64_t)b << 32);
}

static constexpr uint64_t PCG32_DEFAULT_STATE = 0x853c49e6748fea9bULL;   // random.h:40
static constexpr uint64_t PCG32_DEFAULT_STREAM = 0xda3e39cb94b95bdbULL;  // random.h:41
static constexpr uint64_t PCG32_MULT = 0x5851f42d4c957f2dULL;           // random.h:42

// PCG32 (include/util/random.h:45-119)
struct PCG32 {
    uint64_t state, inc;
    PCG32(uint64_t initstate = PCG32_DEFAULT_STATE, uint64_t initseq = PCG32_DEFAULT_STREAM) {
        seed(initstate, initseq);
    }
    void seed(uint64_t initstate, uint64_t initseq) {   // random.h:52-59
        state = 0;
        inc = (initseq << 1) | 1;
        next_u32();
        state += initstate;
        next_u32();
    }
    uint32_t next_u32() {   // random.h:61-68
        const uint64_t old = state;
        state = old * PCG32_MULT + inc;
        const uint32_t xorshifted = (uint32_t)(((old >> 18) ^ old) >> 27);
        const uint32_t rot = (uint32_t)(old >> 59);
        return (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
    }
    // random.h:70-73, C3: low word from the first draw
    uint64_t next_u64() {
        const uint64_t lo = next_u32();
        const uint64_t hi = next_u32();
        return lo | (hi << 32);
    }
    float next_float() {   // random.h:76-81
        uint32_t bits = (next_u32() >> 9) | 0x3f800000u;
        float f;
        std::memcpy(&f, &bits, 4);
        return f - 1.0f;
    }
    double next_double() {   // random.h:84-89
        uint64_t bits = ((uint64_t)next_u32() << 20) | 0x3ff0000000000000ull;
        double d;
        std::memcpy(&d, &bits, 8);
        return d - 1.0;
    }
    V2 next2d() {   // random.h:99-102, C3: x first
        const float x = next_float();
        const float y = next_float();
        return {x, y};
    }
};

// ---------------------------------------------------------------------------------------
// Scene and traversal (stands in for the OptiX GAS/IAS of src/Mesh.cpp, src/Scene.cpp and
// for optixTrace in include/graph/trace.h:25-88)
// ---------------------------------------------------------------------------------------

struct Material {
    V3 diff, spec, emissive;
    float roughness;
    int32_t diffTex = -1, specTex = -1, roughTex = -1;   // MaterialSbtData has*Tex / *Tex
    uint32_t specUsage = 0;                              // SpecTexUsage: 0 Color, 1 OccRoughMetal
};

struct Texture {   // RGBA8, rows in file order (src/Texture.cpp:7-30)
    uint32_t w = 0, h = 0;
    std::vector<uint8_t> rgba;
};

struct EmitterData {   // include/model/Emitter.h:13-55, faces pre-transformed (src/Emitter.cpp:20-27)
    V3 color;
    std::vector<V3> faces;   // 3 per face
};

struct Node {
    float lo[3], hi[3];
    int32_t left, right;     // inner: child indices; leaf: left = first, right = -count
};

struct Scene {
    std::vector<V3> pos, nrm;
    std::vector<float> uv;          // 2 per vertex, or empty (all zero)
    std::vector<Texture> textures;
    std::vector<uint32_t> idx;      // 3 per triangle
    std::vector<uint32_t> triMat;
    std::vector<uint8_t> triVis;    // CX_VIS_* bits per triangle (C9')
    std::vector<Material> mats;
    std::vector<EmitterData> emitters;
    // Evaluation tables of the FP64 adjoint (values and derivatives); `mats` and the emitters'
    // colours drive sampling and every decision.
    std::vector<Material> evalMats;
    std::vector<V3> evalEmit;
    std::vector<Node> nodes;
    std::vector<uint32_t> triOrder;
    float sceneSize = 0.0f;
    V3 bmin, bmax;
    bool bruteForce = false;
    uint32_t numTris() const { return (uint32_t)triMat.size(); }
    V3 vtx(uint32_t tri, int k) const { return pos[idx[tri * 3 + k]]; }
};

// Median-split BVH over triangle centroids, boxes padded outward so that the float box
// test never rejects a box whose triangles the watertight test could report (C5).
static void pad_box(float lo[3], float hi[3]) {
    for (int a = 0; a < 3; a++) {
        const float m = std::max(std::fabs(lo[a]), std::fabs(hi[a]));
        const float e = m * 3.0517578125e-5f + 1e-20f;
        lo[a] -= e;
        hi[a] += e;
    }
}

static int build_node(Scene& s, std::vector<V3>& cent, uint32_t b, uint32_t e, int depth) {
    const int id = (int)s.nodes.size();
    s.nodes.push_back(Node());
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = b; i < e; i++) {
        const uint32_t t = s.triOrder[i];
        for (int k = 0; k < 3; k++) {
            const V3 v = s.vtx(t, k);
            for (int a = 0; a < 3; a++) { lo[a] = std::min(lo[a], v[a]); hi[a] = std::max(hi[a], v[a]); }
        }
        for (int a = 0; a < 3; a++) { clo[a] = std::min(clo[a], cent[t][a]); chi[a] = std::max(chi[a], cent[t][a]); }
    }
    pad_box(lo, hi);
    Node n;
    for (int a = 0; a < 3; a++) { n.lo[a] = lo[a]; n.hi[a] = hi[a]; }
    if (e - b <= 4 || depth > 60) {
        n.left = (int32_t)b;
        n.right = -(int32_t)(e - b);
        s.nodes[id] = n;
        return id;
    }
    int axis = 0;
    float ext = chi[0] - clo[0];
    for (int a = 1; a < 3; a++) if (chi[a] - clo[a] > ext) { ext = chi[a] - clo[a]; axis = a; }
    const uint32_t mid = (b + e) / 2;
    std::nth_element(s.triOrder.begin() + b, s.triOrder.begin() + mid, s.triOrder.begin() + e,
                     [&](uint32_t x, uint32_t y) {
                         if (cent[x][axis] != cent[y][axis]) return cent[x][axis] < cent[y][axis];
                         return x < y;
                     });
    const int l = build_node(s, cent, b, mid, depth + 1);
    const int r = build_node(s, cent, mid, e, depth + 1);
    n.left = l;
    n.right = r;
    s.nodes[id] = n;
    return id;
}

static void build_bvh(Scene& s) {
    const uint32_t nt = s.numTris();
    s.triOrder.resize(nt);
    std::vector<V3> cent(nt);
    for (uint32_t t = 0; t < nt; t++) {
        s.triOrder[t] = t;
        const V3 a = s.vtx(t, 0), b = s.vtx(t, 1), c = s.vtx(t, 2);
        cent[t] = V3((a.x + b.x + c.x) / 3.0f, (a.y + b.y + c.y) / 3.0f, (a.z + b.z + c.z) / 3.0f);
    }
    s.nodes.clear();
    if (nt > 0) build_node(s, cent, 0, nt, 0);
}

struct RayPre {
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

// Woop, Benthin, Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013: per-ray
// permutation and shear.
static inline RayPre ray_pre(const V3& d) {
    const float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    RayPre r;
    r.kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
    r.kx = r.kz + 1; if (r.kx == 3) r.kx = 0;
    r.ky = r.kx + 1; if (r.ky == 3) r.ky = 0;
    if (d[r.kz] < 0.0f) std::swap(r.kx, r.ky);
    r.Sx = d[r.kx] / d[r.kz];
    r.Sy = d[r.ky] / d[r.kz];
    r.Sz = 1.0f / d[r.kz];
    return r;
}

// Returns true and (t, u, v) if the triangle is hit with t in [tmin, tmax]; u weights v1
// and v weights v2 as optixGetTriangleBarycentrics does.
static inline bool tri_hit(const V3& o, const RayPre& r, const V3& v0, const V3& v1, const V3& v2,
                           float tmin, float tmax, float& t, float& u, float& v) {
    const V3 A = v0 - o, B = v1 - o, C = v2 - o;
    const float Ax = A[r.kx] - r.Sx * A[r.kz], Ay = A[r.ky] - r.Sy * A[r.kz];
    const float Bx = B[r.kx] - r.Sx * B[r.kz], By = B[r.ky] - r.Sy * B[r.kz];
    const float Cx = C[r.kx] - r.Sx * C[r.kz], Cy = C[r.ky] - r.Sy * C[r.kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = r.Sz * A[r.kz], Bz = r.Sz * B[r.kz], Cz = r.Sz * C[r.kz];
    const float T = U * Az + V * Bz + W * Cz;
    const float tt = T / det;
    if (!(tt >= tmin && tt <= tmax)) return false;
    t = tt;
    u = V / det;
    v = W / det;
    return true;
}

static inline bool box_visit(const Node& n, const V3& o, const V3& inv, float tmin, float tmax) {
    float tn = tmin, tf = tmax;
    for (int a = 0; a < 3; a++) {
        const float t0 = (n.lo[a] - o[a]) * inv[a];
        const float t1 = (n.hi[a] - o[a]) * inv[a];
        if (std::isnan(t0) || std::isnan(t1)) return true;
        tn = std::max(tn, std::min(t0, t1));
        tf = std::min(tf, std::max(t0, t1));
    }
    return tn <= tf * 1.00002f + 1e-30f;
}

struct Hit { uint32_t prim; float t, u, v; };

static Hit trace_closest(const Scene& s, const V3& o, const V3& d, float tmin, float tmax) {
    Hit best{INVALID_PRIM, tmax, 0.0f, 0.0f};
    const RayPre rp = ray_pre(d);
    auto consider = [&](uint32_t tri) {
        if (!(s.triVis[tri] & CX_VIS_SURFACE)) return;
        float t, u, v;
        if (!tri_hit(o, rp, s.vtx(tri, 0), s.vtx(tri, 1), s.vtx(tri, 2), tmin, tmax, t, u, v)) return;
        if (best.prim == INVALID_PRIM || t < best.t || (t == best.t && tri < best.prim)) best = {tri, t, u, v};
    };
    if (s.bruteForce || s.nodes.empty()) {
        for (uint32_t t = 0; t < s.numTris(); t++) consider(t);
        return best;
    }
    const V3 inv(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const Node& n = s.nodes[stack[--sp]];
        const float tcap = (best.prim == INVALID_PRIM) ? tmax : best.t;
        if (!box_visit(n, o, inv, tmin, tcap)) continue;
        if (n.right < 0) {
            for (int i = 0; i < -n.right; i++) consider(s.triOrder[n.left + i]);
        } else {
            stack[sp++] = n.right;
            stack[sp++] = n.left;
        }
    }
    return best;
}

static bool trace_occluded(const Scene& s, const V3& o, const V3& d, float tmin, float tmax) {
    const RayPre rp = ray_pre(d);
    float t, u, v;
    auto blocks = [&](uint32_t tri) {
        return (s.triVis[tri] & CX_VIS_OCCLUSION) &&
               tri_hit(o, rp, s.vtx(tri, 0), s.vtx(tri, 1), s.vtx(tri, 2), tmin, tmax, t, u, v);
    };
    if (s.bruteForce || s.nodes.empty()) {
        for (uint32_t tri = 0; tri < s.numTris(); tri++)
            if (blocks(tri)) return true;
        return false;
    }
    const V3 inv(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const Node& n = s.nodes[stack[--sp]];
        if (!box_visit(n, o, inv, tmin, tmax)) continue;
        if (n.right < 0) {
            for (int i = 0; i < -n.right; i++)
                if (blocks(s.triOrder[n.left + i])) return true;
        } else {
            stack[sp++] = n.right;
            stack[sp++] = n.left;
        }
    }
    return false;
}

// ---------------------------------------------------------------------------------------
// Interactions (include/graph/Interaction.h)
// ---------------------------------------------------------------------------------------

// PackedInteraction (Interaction.h:59-86): HitInfo {mesh, instanceId, primitiveIdx, bary}
// + inRayDir + valid.  One flattened mesh, instance 0 (C8).
struct Packed {
    uint32_t prim = 0;
    float bu = 0.0f, bv = 0.0f;
    V3 inDir;
    bool valid = false;
    bool hasMesh = false;   // meshSbtData != nullptr
};

// Interaction (Interaction.h:88-184) without tangent and uv, which no untextured shading
// path reads.
struct Inter {
    uint32_t prim = 0;
    V3 pos, inDir, normal;
    float bu = 0.0f, bv = 0.0f;
    bool valid = false;
    bool hasMesh = false;
    Packed pack() const {   // Interaction.h:174-178
        Packed p;
        p.prim = prim; p.bu = bu; p.bv = bv; p.inDir = inDir; p.valid = valid; p.hasMesh = hasMesh;
        return p;
    }
};

// InterpolateBary (Interaction.h:18-25): (1-u-v)*v0 + u*v1 + v*v2
static inline V3 interp(float u, float v, const V3& a, const V3& b, const V3& c) {
    return ((1.0f - u - v) * a + u * b) + v * c;
}

// Rehydration constructor (Interaction.h:115-172); normals stay un-normalised (C7).
static Inter rehydrate(const Scene& s, const Packed& p) {
    Inter r;
    r.valid = p.valid; r.inDir = p.inDir; r.prim = p.prim; r.bu = p.bu; r.bv = p.bv; r.hasMesh = p.hasMesh;
    if (!r.valid) { r.valid = false; return r; }
    if (r.prim >= s.numTris()) { r.valid = false; return r; }
    const uint32_t i0 = s.idx[r.prim * 3 + 0], i1 = s.idx[r.prim * 3 + 1], i2 = s.idx[r.prim * 3 + 2];
    r.pos = interp(r.bu, r.bv, s.pos[i0], s.pos[i1], s.pos[i2]);
    r.normal = interp(r.bu, r.bv, s.nrm[i0], s.nrm[i1], s.nrm[i2]);
    return r;
}


// Ray counters, one cache line per OpenMP thread, summed after each pass.
struct alignas(64) RayCount {
    uint64_t closest = 0, shadow = 0;
};
static RayCount g_rays[256];
static inline RayCount& ray_count() {
#ifdef _OPENMP
    return g_rays[omp_get_thread_num() & 255];
#else
    return g_rays[0];
#endif
}
static void take_rays(uint64_t& closest, uint64_t& shadow) {
    closest = shadow = 0;
    for (RayCount& r : g_rays) {
        closest += r.closest;
        shadow += r.shadow;
        r.closest = r.shadow = 0;
    }
}

// TraceWithDataPointer<Interaction> + __closesthit__mesh / __miss__main
// (include/graph/trace.h:25-45, shader/mesh_hit.cu:17-69, shader/miss.cu:11-19).
static void trace_surface(const Scene& s, const V3& o, const V3& d, float tmin, float tmax, Inter& io) {
    ray_count().closest++;
    const Hit h = trace_closest(s, o, d, tmin, tmax);
    if (h.prim == INVALID_PRIM) {   // miss.cu:11-19
        io.valid = false;
        io.inDir = d;
        return;
    }
    io.hasMesh = true;
    io.inDir = d;
    io.valid = true;
    io.prim = h.prim;
    io.bu = h.u; io.bv = h.v;
    const uint32_t i0 = s.idx[h.prim * 3 + 0], i1 = s.idx[h.prim * 3 + 1], i2 = s.idx[h.prim * 3 + 2];
    io.pos = interp(h.u, h.v, s.pos[i0], s.pos[i1], s.pos[i2]);
    io.normal = normalize(interp(h.u, h.v, s.nrm[i0], s.nrm[i1], s.nrm[i2]));
}

// TraceOcclusion (include/graph/trace.h:63-88); occluded = any hit in [tmin, tmax].
static bool trace_occ(const Scene& s, const V3& o, const V3& d, float tmin, float tmax) {
    ray_count().shadow++;
    return trace_occluded(s, o, d, tmin, tmax);
}

// ---------------------------------------------------------------------------------------
// BSDF (shader/brdf.cu)
// ---------------------------------------------------------------------------------------

struct BrdfEval { V3 brdfResult; float samplingPdf = 0.0f; V3 emission; float roughness = 0.0f; };
struct BrdfSample { V3 outDir, brdfVal; float samplingPdf = 0.0f; float roughness = 0.0f; };

static inline float D_GGX(float NdotH, float roughness) {   // brdf.cu:99-104
    const float a2 = roughness * roughness;
    const float d = (NdotH * a2 - NdotH) * NdotH + 1.0f;
    return a2 / (PI * d * d);
}
static inline float V_SmithJointGGX(float NdotL, float NdotV, float roughness) {   // brdf.cu:106-112
    const float a2 = roughness * roughness;
    const float lambdaV = NdotL * std::sqrt(NdotV * NdotV * (1 - a2) + a2);
    const float lambdaL = NdotV * std::sqrt(NdotL * NdotL * (1 - a2) + a2);
    return 0.5f / (lambdaV + lambdaL);
}
static inline V3 FresnelSchlick(const V3& F0, float VdotH) {   // brdf.cu:34-37
    const float p = pow5(gmax(0.0f, 1.0f - VdotH));
    return F0 + (V3(1.0f) - F0) * p;
}

// tex2D<float4> of a texture made by Texture::LoadToDevice (src/Texture.cpp:44-76): wrap
// addressing, linear filtering, normalized coordinates, normalized-float reads.  The filter is
// the CUDA programming guide's (texture fetching, linear filtering): with x = frac(u) * w - 0.5,
// i = floor(x), alpha = frac(x) held with 8 fractional bits,
//   T = (1-alpha)(1-beta) T[i,j] + alpha(1-beta) T[i+1,j] + (1-alpha)beta T[i,j+1] + alpha beta T[i+1,j+1],
// indices wrapped modulo the size.  Canonical here: alpha and beta are rounded to the nearest
// multiple of 1/256, and the weighted sum, an integer in 1/(65536*255) units, is rounded once.
// texel_sum returns that integer per channel; a non-finite coordinate reads as 0.
static void texel_sum(const Texture& t, float u, float v, uint32_t sum[3]) {
    if (!std::isfinite(u)) u = 0.0f;
    if (!std::isfinite(v)) v = 0.0f;
    const float x = (u - std::floor(u)) * (float)t.w - 0.5f;
    const float y = (v - std::floor(v)) * (float)t.h - 0.5f;
    const float fx = std::floor(x), fy = std::floor(y);
    const int64_t ia = (int64_t)std::nearbyint((x - fx) * 256.0f), ib = (int64_t)std::nearbyint((y - fy) * 256.0f);
    const int64_t W = t.w, H = t.h;
    const int64_t xi[2] = {(((int64_t)fx % W) + W) % W, (((int64_t)fx + 1) % W + W) % W};
    const int64_t yi[2] = {(((int64_t)fy % H) + H) % H, (((int64_t)fy + 1) % H + H) % H};
    const int64_t wx[2] = {256 - ia, ia}, wy[2] = {256 - ib, ib};
    for (int c = 0; c < 3; c++) {
        int64_t acc = 0;
        for (int b = 0; b < 2; b++)
            for (int a = 0; a < 2; a++) acc += wx[a] * wy[b] * t.rgba[(size_t)(yi[b] * W + xi[a]) * 4 + c];
        sum[c] = (uint32_t)acc;
    }
}

static V3 tex_fetch(const Texture& t, float u, float v) {
    uint32_t q[3];
    texel_sum(t, u, v, q);
    return V3((float)q[0] / 16711680.0f, (float)q[1] / 16711680.0f, (float)q[2] / 16711680.0f);
}

// The interaction's uv: InterpolateBary of the vertex uvs (Interaction.h:171).
static void inter_uv(const Scene& s, uint32_t prim, float bu, float bv, float& u, float& v) {
    u = v = 0.0f;
    if (s.uv.empty()) return;
    const uint32_t i0 = s.idx[prim * 3], i1 = s.idx[prim * 3 + 1], i2 = s.idx[prim * 3 + 2];
    u = ((1.0f - bu - bv) * s.uv[2 * i0] + bu * s.uv[2 * i1]) + bv * s.uv[2 * i2];
    v = ((1.0f - bu - bv) * s.uv[2 * i0 + 1] + bu * s.uv[2 * i1 + 1]) + bv * s.uv[2 * i2 + 1];
}

// Material fetch of ggx_eval / ggx_sample (brdf.cu:121-138): GetFromTexIfPossible for the
// diffuse colour, specular value and roughness, the OccRoughMetal channel mapping, and
// specF0 = mix(diffColor, specF0, metal).
static inline void material_params(const Scene& s, const Inter& it, V3& diffColor, V3& specF0, float& roughness,
                                   V3& emissive) {
    const Material& m = s.mats[s.triMat[it.prim]];
    float u = 0.0f, v = 0.0f;
    if (m.diffTex >= 0 || m.specTex >= 0 || m.roughTex >= 0) inter_uv(s, it.prim, it.bu, it.bv, u, v);
    diffColor = m.diffTex >= 0 ? tex_fetch(s.textures[m.diffTex], u, v) : m.diff;
    specF0 = m.specTex >= 0 ? tex_fetch(s.textures[m.specTex], u, v) : m.spec;
    roughness = m.roughTex >= 0 ? tex_fetch(s.textures[m.roughTex], u, v).x : m.roughness;
    float metal = 0.0f;
    if (m.specUsage == 1) {
        roughness = specF0.y;
        metal = specF0.z;
        specF0 = V3(1.0f);
    }
    specF0 = mix(diffColor, specF0, metal);
    emissive = m.emissive;
}

// __direct_callable__ggx_eval (brdf.cu:115-190)
static BrdfEval ggx_eval(const Scene& s, const Inter& it, const V3& outDir) {
    V3 diffColor, specF0, emissive;
    float roughness;
    material_params(s, it, diffColor, specF0, roughness, emissive);
    BrdfEval r;
    const V3 lightDir = outDir;
    const V3 viewDir = -it.inDir;
    const V3 normal = it.normal;
    const float nDotV = dot(normal, viewDir);
    const float nDotL = dot(normal, lightDir);
    if (nDotV <= 0.0f || nDotL < 0.0f) {
        r.brdfResult = V3(0.0f); r.emission = V3(0.0f); r.samplingPdf = 0.0f; r.roughness = 0.0f;
        return r;
    }
    const V3 diffBrdf = diffColor / PI;
    V3 specBrdf(0.0f);
    const float diffProb = dot(diffColor, V3(1.0f)) / (dot(diffColor, V3(1.0f)) + dot(specF0, V3(1.0f)));
    const V3 halfway = normalize(lightDir + viewDir);
    const float nDotH = dot(halfway, normal);
    const float lDotH = dot(halfway, lightDir);
    if ((double)dot(specF0, specF0) > 1e-6) {
        const float d = D_GGX(nDotH, roughness);
        const float v = V_SmithJointGGX(nDotL, nDotV, roughness);
        const V3 f = FresnelSchlick(specF0, lDotH);
        specBrdf = d * v * f;
    }
    r.brdfResult = (diffBrdf + specBrdf) * nDotL;
    r.samplingPdf = diffProb * nDotL / PI + (1.0f - diffProb) * D_GGX(nDotH, roughness) * nDotH / (4.0f * lDotH);
    r.emission = emissive;
    r.roughness = roughness;
    return r;
}

static inline M3 local_frame(const V3& localZ) {   // brdf.cu:192-212
    const float x = localZ.x, y = localZ.y, z = localZ.z;
    const float sz = (z >= 0) ? 1.0f : -1.0f;
    const float a = 1 / (sz + z);
    const float ya = y * a;
    const float b = x * ya;
    const float c = x * sz;
    M3 m;
    m.c0 = V3(c * x * a - 1, sz * b, c);
    m.c1 = V3(b, y * ya - sz, y);
    m.c2 = localZ;
    return m;
}

// __direct_callable__ggx_sample (brdf.cu:214-338)
static BrdfSample ggx_sample(const Scene& s, const Inter& it, PCG32& rng) {
    V3 diffColor, specF0, emissive;
    float roughness;
    material_params(s, it, diffColor, specF0, roughness, emissive);
    const V3 normal = it.normal;
    const V3 viewDir = -it.inDir;
    BrdfSample r;
    r.roughness = roughness;
    const float nDotV = dot(normal, viewDir);
    if (nDotV <= 0) {
        r.outDir = V3(0.0f); r.brdfVal = V3(0.0f); r.samplingPdf = 0.0f;
        return r;
    }
    const M3 frame = local_frame(normal);
    const float diffProb = dot(diffColor, V3(1.0f)) / (dot(diffColor, V3(1.0f)) + dot(specF0, V3(1.0f)));
    if (rng.next_float() < diffProb) {
        const float phi = rng.next_float() * 2.0f * PI;
        const float rr = std::sqrt(rng.next_float());
        float sp, cp;
        sincos_cw(phi, &sp, &cp);
        const float x = rr * cp;
        const float y = rr * sp;
        const float z = std::sqrt(gclamp(1.0f - x * x - y * y, 0.0f, 1.0f));
        r.outDir = mul(frame, V3(x, y, z));
    } else {
        const float u = rng.next_float();
        const float mz = std::sqrt(gclamp((1 - u) / (1 + (roughness * roughness - 1) * u), 0.0f, 1.0f));
        const float mr = std::sqrt(gclamp(1 - mz * mz, 0.0f, 1.0f));
        const float phi = rng.next_float() * 2.0f * PI;
        float sp, cp;
        sincos_cw(phi, &sp, &cp);
        const V3 m(mr * cp, mr * sp, mz);
        r.outDir = reflect(it.inDir, mul(frame, m));
        if (dot(r.outDir, normal) < 0.0f) {
            r.outDir = V3(0.0f); r.brdfVal = V3(0.0f); r.samplingPdf = 0.0f;
            return r;
        }
    }
    const V3 diffBrdf = diffColor / PI;
    const V3 lightDir = r.outDir;
    const float nDotL = dot(normal, lightDir);
    const V3 halfway = normalize(lightDir + viewDir);
    const float nDotH = dot(halfway, normal);
    const float lDotH = dot(halfway, lightDir);
    V3 specBrdf(0.0f);
    if ((double)dot(specF0, specF0) > 1e-6) {
        const float d = D_GGX(nDotH, roughness);
        const float v = V_SmithJointGGX(nDotL, nDotV, roughness);
        const V3 f = FresnelSchlick(specF0, lDotH);
        specBrdf = d * v * f;
    }
    r.outDir = lightDir;
    r.samplingPdf = diffProb * nDotL / PI + (1.0f - diffProb) * D_GGX(nDotH, roughness) * nDotH / (4.0f * lDotH);
    r.brdfVal = (diffBrdf + specBrdf) * nDotL;
    return r;
}

// ---------------------------------------------------------------------------------------
// Emitters (include/graph/sample_emitter.h:3-12, include/model/Emitter.h:21-54)
// ---------------------------------------------------------------------------------------

struct EmitterSample { V3 pos, color; float p; };

static EmitterSample sample_emitter(const Scene& s, PCG32& rng) {
    const uint64_t count = s.emitters.size();
    const uint64_t ei = rng.next_u64() % count;
    const EmitterData& e = s.emitters[ei];
    const uint64_t faceIdx = rng.next_u64() % (uint64_t)(e.faces.size() / 3);   // SampleFaceAreaWeighted
    const V3 v0 = e.faces[faceIdx * 3 + 0], v1 = e.faces[faceIdx * 3 + 1], v2 = e.faces[faceIdx * 3 + 2];
    float r0 = rng.next_float();
    float r1 = rng.next_float();
    float r2 = rng.next_float();
    const float rSum = r0 + r1 + r2;
    r0 /= rSum;
    r1 /= rSum;
    r2 /= rSum;
    const V3 p = r0 * v0 + r1 * v1 + r2 * v2;
    return {p, e.color, 1.0f};
}

// ---------------------------------------------------------------------------------------
// Path records (include/graph/restir/PrefixPath.h, SuffixPath.h, Reservoir.h, RestirGBuffer.h)
// ---------------------------------------------------------------------------------------

struct Prefix {   // PrefixPath.h:7-132
    PCG32 rng;
    Packed primaryInt, reconInt, lastInt;
    uint32_t flags = 0;
    V3 f, postReconF;
    float p = 0.0f, pathLen = 0.0f;
    V3 reconOutDir;
    uint32_t length() const { return flags & 0xFFu; }
    void setLength(uint32_t l) { flags &= ~0xFFu; flags |= l & 0xFFu; }
    uint32_t reconIdx() const { return (flags >> 8u) & 0xFFu; }
    void setReconIdx(uint32_t r) { flags &= ~0xFF00u; flags |= (r & 0xFFu) << 8u; }
    bool valid() const { return (flags & (1u << 16u)) != 0; }
    void setValid(bool v) { flags &= ~(1u << 16u); if (v) flags |= 1u << 16u; }
    bool nee() const { return (flags & (1u << 17u)) != 0; }
    void setNee(bool v) { flags &= ~(1u << 17u); if (v) flags |= 1u << 17u; }
    void reset() { flags = 0u; pathLen = 0.0f; }
};

// Shift constructor (PrefixPath.h:60-81).  The primitive check of :76-80 cannot fire in a
// single flattened mesh, where every stored primitive index is in range.
static Prefix prefix_shifted(const Scene& s, const Prefix& other, const V3& f, const Packed& primaryInt) {
    Prefix r = other;
    r.primaryInt = primaryInt;
    r.f = f;
    if (!r.valid() || !r.reconInt.hasMesh) return r;
    if (s.numTris() <= r.reconInt.prim) r.setValid(false);
    return r;
}

struct Suffix {   // SuffixPath.h:9-104
    PCG32 rng;
    Packed reconInt, lastPrefixInt;
    uint32_t flags = 0;
    V3 f, postReconF;
    float p = 0.0f;
    V3 reconOutDir;
    uint32_t length() const { return flags & 0xFFu; }
    void setLength(uint32_t l) { flags &= ~0xFFu; flags |= l & 0xFFu; }
    uint32_t reconIdx() const { return (flags >> 8u) & 0xFFu; }
    void setReconIdx(uint32_t r) { flags &= ~0xFF00u; flags |= (r & 0xFFu) << 8u; }
    bool valid() const { return (flags & (1u << 16u)) != 0; }
    void setValid(bool v) { flags &= ~(1u << 16u); if (v) flags |= 1u << 16u; }
    void reset() { flags = 0u; }
};

static Suffix suffix_shifted(const Suffix& other, const Packed& lastPrefixInt, const V3& f) {   // SuffixPath.h:51-65
    Suffix r = other;
    r.lastPrefixInt = lastPrefixInt;
    r.f = f;
    return r;
}

template <typename T>
struct Reservoir {   // Reservoir.h:8-51
    T sample;
    float wSum = 0.0f, confidence = 0.0f;
    void reset() { wSum = 0.0f; confidence = 0.0f; }
    bool update(const T& s, float risWeight, PCG32& rng, float inputConfidence) {
        if (std::isnan(risWeight) || std::isinf(risWeight)) risWeight = 0.0f;
        wSum += risWeight;
        confidence += inputConfidence;
        confidence = gmin(confidence, CONFIDENCE_MAX);
        if (rng.next_float() < risWeight / wSum) {
            sample = s;
            return true;
        }
        return false;
    }
    void finalize() {
        const float pHat = luminance(sample.f);
        if (pHat == 0.0f) wSum = 0.0f;
        else wSum /= pHat;
    }
};

struct GBuf {   // RestirGBuffer.h:5-28
    PCG32 rng;
    uint32_t prevX = 0, prevY = 0;
    bool primaryIntValid = false;
};

struct Neighbor { uint32_t pixelIdx = 0; float distance = 0.0f; };   // PrefixAccelStruct.h:8-27

// ---------------------------------------------------------------------------------------
// Pairwise MIS, Jacobian, neighbour selection (include/graph/restir/path_gen.h:12-72)
// ---------------------------------------------------------------------------------------

static inline float canonical_pairwise_mis(const V3& atBasis, const V3& atNeighbor, float J, float K, float canonM,
                                           float neighM) {   // path_gen.h:15-31
    float w = 1.0f;
    if (luminance(atBasis) > 0.0f) {
        const float atBasisTerm = luminance(atBasis) * canonM;
        w = atBasisTerm / (atBasisTerm + luminance(atNeighbor) * J * neighM * K);
    }
    return w;
}

static inline float neighbor_pairwise_mis(const V3& atBasis, const V3& atNeighbor, float J, float K, float canonM,
                                          float neighM) {   // path_gen.h:33-48
    float w = 0.0f;
    if (luminance(atNeighbor) > 0.0f) {
        w = luminance(atNeighbor) * neighM /
            (luminance(atNeighbor) * neighM + luminance(atBasis) * J * canonM / K);
    }
    return w;
}

static inline void select_spatial_neighbor(uint32_t px, uint32_t py, uint32_t R, PCG32& rng, uint32_t& nx,
                                           uint32_t& ny) {   // path_gen.h:50-57
    const uint32_t W = 2 * R + 1;
    const uint32_t x = (px + R) - (rng.next_u32() % W);
    const uint32_t y = (py + R) - (rng.next_u32() % W);
    if (x == px && y == py) { nx = px + 1; ny = py + 1; return; }
    nx = x; ny = y;
}

static inline float reconnection_jacobian(const V3& oldXi, const V3& newXi, const V3& targetPos,
                                          const V3& targetNormal) {   // path_gen.h:59-72
    const float term1 = std::fabs(dot(normalize(newXi - targetPos), targetNormal) /
                                  dot(normalize(oldXi - targetPos), targetNormal));
    const float term2 = dot(targetPos - oldXi, targetPos - oldXi);
    const float term3 = dot(targetPos - newXi, targetPos - newXi);
    const float result = term1 * term2 / term3;
    if (std::isinf(result) || std::isnan(result)) return 0.0f;
    return gmax(0.0f, result);
}

// ---------------------------------------------------------------------------------------
// Frame parameters
// ---------------------------------------------------------------------------------------

struct Params {
    cx_launch_params lp;
    uint8_t front = 0, back = 0;
    V3 camPos, camU, camV, camW;
};

static inline bool is_restir_pt(const Params& P) { return P.lp.rendererType == CX_RESTIR_PT; }
static inline bool is_cond(const Params& P) { return P.lp.rendererType == CX_CONDITIONAL_RESTIR; }

// ---------------------------------------------------------------------------------------
// Path generation (include/graph/restir/path_gen.h:74-531)
// ---------------------------------------------------------------------------------------

static void trace_prefix(const Scene& s, Prefix& prefix, const V3& origin, const V3& dir, PCG32& rng,
                         const Params& P) {   // path_gen.h:74-250
    V3 currentPos = origin;
    V3 currentDir = dir;
    prefix.reset();
    prefix.rng = rng;
    prefix.p = 1.0f;
    prefix.f = V3(1.0f);
    prefix.postReconF = V3(1.0f);
    prefix.setValid(true);
    Inter interaction;
    const uint32_t maxLen = is_restir_pt(P) ? (uint32_t)P.lp.maxPathLen : (uint32_t)P.lp.prefixLen;
    bool postRecon = false;
    for (uint32_t traceIdx = 0; traceIdx < maxLen; ++traceIdx) {
        trace_surface(s, currentPos, currentDir, 1e-3f, 1e16f, interaction);
        if (!interaction.valid) { prefix.setValid(false); return; }
        prefix.pathLen += distance(currentPos, interaction.pos);
        prefix.setLength(prefix.length() + 1);
        if (prefix.length() == 1) prefix.primaryInt = interaction.pack();
        if (is_cond(P) && traceIdx == maxLen - 1) {   // :130-142
            if (!postRecon) {
                prefix.reconInt = interaction.pack();
                prefix.setReconIdx(prefix.length());
                prefix.reconOutDir = V3(0.0f);
            }
            prefix.setValid(true);
            prefix.lastInt = interaction.pack();
            return;
        }
        const bool forceNee = is_restir_pt(P) && traceIdx == maxLen - 1;
        if (rng.next_float() < P.lp.neeProb || forceNee) {   // :144-211
            prefix.p *= P.lp.neeProb;
            const EmitterSample es = sample_emitter(s, rng);
            const V3 lightDir = normalize(es.pos - interaction.pos);
            const float dist = length(es.pos - interaction.pos);
            if (trace_occ(s, interaction.pos, lightDir, 1e-3f, dist)) { prefix.setValid(false); return; }
            if (!postRecon) {   // :165-187, re-traces the same ray
                trace_surface(s, currentPos, currentDir, 1e-3f, 1e16f, interaction);
                if (interaction.valid) {
                    prefix.reconInt = interaction.pack();
                    prefix.setReconIdx(prefix.length());
                    prefix.reconOutDir = currentDir;
                    postRecon = true;
                } else {
                    prefix.setValid(false);
                    return;
                }
            }
            const BrdfEval be = ggx_eval(s, interaction, lightDir);
            if (be.samplingPdf <= 0.0f) { prefix.setValid(false); return; }
            const float distFactor = 1.0f / (dist * dist);
            prefix.f *= be.brdfResult * es.color * distFactor;
            if (postRecon) prefix.postReconF *= be.brdfResult * es.color * distFactor;
            if (prefix.length() == 1) prefix.f += be.emission * distFactor;
            prefix.setNee(true);
            prefix.setValid(true);
            prefix.lastInt = interaction.pack();
            return;
        }
        const BrdfSample bs = ggx_sample(s, interaction, rng);   // :213-222
        if (bs.samplingPdf <= 0.0f) { prefix.setValid(false); return; }
        const float reconDistance = distance(currentPos, interaction.pos);
        const float reconRoughness = bs.roughness;
        const bool intCanRecon = reconDistance > P.lp.reconMinDistance && reconRoughness > P.lp.reconMinRoughness &&
                                 prefix.length() >= 2;
        if (!postRecon && intCanRecon) {
            prefix.reconInt = interaction.pack();
            prefix.setReconIdx(prefix.length());
            prefix.reconOutDir = bs.outDir;
            postRecon = true;
        }
        currentPos = interaction.pos;
        currentDir = bs.outDir;
        prefix.f *= bs.brdfVal;
        prefix.p *= bs.samplingPdf * (1.0f - P.lp.neeProb);
        if (postRecon) prefix.postReconF *= bs.brdfVal;
    }
    prefix.lastInt = interaction.pack();
}

// reconVertex (a test diagnostic, not part of the reference's path): receives the suffix vertex
// at which a BSDF-sampled reconnection was chosen (its length at that point), 0 otherwise.
static void trace_suffix(const Scene& s, Suffix& suffix, const Prefix& prefix, PCG32& rng,
                         const Params& P, uint32_t* reconVertex = nullptr) {   // path_gen.h:252-387
    if (reconVertex) *reconVertex = 0;
    suffix.f = V3(1.0f);
    suffix.p = 1.0f;
    suffix.postReconF = V3(1.0f);
    suffix.rng = rng;
    suffix.lastPrefixInt = prefix.lastInt;
    suffix.reset();
    if (!prefix.valid() || prefix.nee()) return;
    const uint32_t maxLen = (uint32_t)P.lp.maxPathLen - prefix.length();
    Inter interaction = rehydrate(s, prefix.lastInt);
    if (!interaction.valid) return;
    bool postRecon = false;
    V3 lastDir(0.0f);
    for (uint32_t vertIdx = 0; vertIdx < maxLen; ++vertIdx) {
        const V3 lastPos = interaction.pos;
        if (suffix.length() > 0) {
            trace_surface(s, lastPos, lastDir, 1e-3f, 1e16f, interaction);
            if (!interaction.valid) return;
        }
        const bool forceNee = vertIdx == maxLen - 1;
        if (rng.next_float() < P.lp.neeProb || forceNee) {
            if (!forceNee) suffix.p *= P.lp.neeProb;
            const EmitterSample em = sample_emitter(s, rng);
            const V3 lightVec = em.pos - interaction.pos;
            const float lightLen = length(lightVec);
            const V3 lightDir = normalize(lightVec);
            if (any_nan_inf(lightDir)) return;
            if (trace_occ(s, interaction.pos, lightDir, 1e-3f, lightLen)) return;
            const BrdfEval be = ggx_eval(s, interaction, lightDir);
            if (be.samplingPdf <= 0.0f) return;
            suffix.setLength(suffix.length() + 1);
            if (!postRecon) {
                Inter reconInt;
                trace_surface(s, interaction.pos, lightDir, 1e-3f, 1e16f, reconInt);
                if (!reconInt.valid) return;
                suffix.reconInt = reconInt.pack();
                suffix.setReconIdx(suffix.length());
                suffix.reconOutDir = V3(0.0f);
                postRecon = true;
            }
            const float distFactor = 1.0f / (lightLen * lightLen);
            suffix.f *= be.brdfResult * em.color * distFactor;
            suffix.postReconF *= em.color * distFactor;
            if (suffix.reconIdx() < suffix.length()) suffix.postReconF *= be.brdfResult;
            suffix.setValid(true);
            return;
        }
        const BrdfSample bs = ggx_sample(s, interaction, rng);
        if (bs.samplingPdf <= 0.0f) return;
        suffix.p *= bs.samplingPdf * (1.0f - P.lp.neeProb);
        suffix.f *= bs.brdfVal;
        if (postRecon) suffix.postReconF *= bs.brdfVal;
        const float reconDistance = distance(lastPos, interaction.pos);
        const float reconRoughness = bs.roughness;
        const bool intCanRecon = reconDistance > P.lp.reconMinDistance && reconRoughness > P.lp.reconMinRoughness &&
                                 prefix.length() >= 1;
        if (!postRecon && intCanRecon) {
            suffix.reconInt = interaction.pack();
            suffix.setReconIdx(prefix.length());   // :378, reference quirk kept
            suffix.reconOutDir = bs.outDir;
            postRecon = true;
            if (reconVertex) *reconVertex = suffix.length();
        }
        lastDir = bs.outDir;
        suffix.setLength(suffix.length() + 1);
    }
}

static V3 trace_complete_path(const Scene& s, const V3& origin, const V3& dir, PCG32& rng,
                              const Params& P) {   // path_gen.h:389-472
    V3 currentPos = origin;
    V3 currentDir = dir;
    uint64_t currentDepth = 0;
    V3 radiance(0.0f);
    V3 throughput(1.0f);
    for (uint32_t traceIdx = 0; traceIdx < (uint32_t)P.lp.maxPathLen; ++traceIdx) {
        Inter interaction;
        trace_surface(s, currentPos, currentDir, 1e-3f, 1e16f, interaction);
        if (!interaction.valid) break;
        ++currentDepth;
        if (rng.next_float() < P.lp.neeProb) {
            throughput /= P.lp.neeProb;
            const EmitterSample es = sample_emitter(s, rng);
            const V3 lightDir = normalize(es.pos - interaction.pos);
            const float dist = length(es.pos - interaction.pos);
            if (trace_occ(s, interaction.pos, lightDir, 1e-3f, dist)) return V3(0.0f);
            const BrdfEval be = ggx_eval(s, interaction, lightDir);
            const float distFactor = 1.0f / (dist * dist);
            radiance = throughput * be.brdfResult * es.color * distFactor;
            if (currentDepth == 1) radiance += be.emission;
            break;
        }
        const BrdfSample bs = ggx_sample(s, interaction, rng);
        if (bs.samplingPdf <= 0.0f) break;
        currentPos = interaction.pos;
        currentDir = bs.outDir;
        throughput *= bs.brdfVal / bs.samplingPdf;
        throughput /= (1.0f - P.lp.neeProb);
    }
    return radiance;
}

static bool trace_prefix_for_final_gather(const Scene& s, V3& throughput, float& p, Inter& lastInt,
                                          const V3& origin, const V3& dir, uint64_t maxLen, PCG32& rng,
                                          const Params& P) {   // path_gen.h:474-531
    (void)P;
    V3 currentPos = origin;
    V3 currentDir = dir;
    throughput = V3(1.0f);
    p = 1.0f;
    lastInt.valid = false;
    for (uint32_t traceIdx = 0; traceIdx < maxLen; ++traceIdx) {
        trace_surface(s, currentPos, currentDir, 1e-3f, 1e16f, lastInt);
        if (!lastInt.valid) return false;
        const BrdfSample bs = ggx_sample(s, lastInt, rng);
        if (bs.samplingPdf <= 0.0f) return false;
        currentPos = lastInt.pos;
        currentDir = bs.outDir;
        throughput *= bs.brdfVal;
        p *= bs.samplingPdf;
    }
    return true;
}

// ---------------------------------------------------------------------------------------
// Hybrid shifts (include/graph/restir/prefix_reuse.h:10-99, suffix_reuse.h:10-123)
// ---------------------------------------------------------------------------------------

static V3 shift_prefix(const Scene& s, const Prefix& src, const Prefix& dst, float& jacobian) {
    jacobian = 0.0f;
    if (!src.valid() || !src.valid() || src.reconIdx() < 2) return V3(0.0f);   // prefix_reuse.h:20
    const Inter srcPrimaryInt = rehydrate(s, src.primaryInt);
    if (!srcPrimaryInt.valid) return V3(0.0f);
    Inter dstPrimaryInt = rehydrate(s, dst.primaryInt);
    if (!dstPrimaryInt.valid) return V3(0.0f);
    Inter srcReconInt = rehydrate(s, src.reconInt);
    if (!srcReconInt.valid) return V3(0.0f);
    const uint32_t reconVertCount = (uint32_t)std::max<int>((int)(src.reconIdx() - 2), 0);
    Inter& currInt = dstPrimaryInt;
    PCG32 srcRng = src.rng;
    V3 throughput(1.0f);
    for (uint32_t idx = 0; idx < reconVertCount; ++idx) {
        srcRng.next_float();   // :44 dummy draw for the NEE decision
        const BrdfSample bs = ggx_sample(s, currInt, srcRng);
        if (bs.samplingPdf <= 0.0f) return V3(0.0f);
        throughput *= bs.brdfVal;
        trace_surface(s, currInt.pos, bs.outDir, 1e-3f, 1e16f, currInt);
        if (!currInt.valid) return V3(0.0f);
    }
    const V3 reconVec = srcReconInt.pos - currInt.pos;
    const float reconLen = length(reconVec);
    const V3 reconDir = normalize(reconVec);
    if (std::isinf(reconLen) || any_nan_inf(reconDir)) return V3(0.0f);
    if (trace_occ(s, currInt.pos, reconDir, 1e-3f, reconLen)) return V3(0.0f);
    const BrdfEval b1 = ggx_eval(s, currInt, reconDir);
    if (b1.samplingPdf <= 0.0f) return V3(0.0f);
    throughput *= b1.brdfResult;
    srcReconInt.inDir = reconDir;
    const BrdfEval b2 = ggx_eval(s, srcReconInt, src.reconOutDir);
    if (b2.samplingPdf <= 0.0f) return V3(0.0f);
    throughput *= b2.brdfResult;
    jacobian = reconnection_jacobian(srcPrimaryInt.pos, currInt.pos, srcReconInt.pos, srcReconInt.normal);
    return gmax(0.0f, throughput * src.postReconF);
}

static V3 shift_suffix(const Scene& s, const Suffix& src, const Inter& dstLastPrefixInt,
                       float& jacobian) {   // suffix_reuse.h:10-108
    jacobian = 0.0f;
    if (!src.valid()) return V3(0.0f);
    if (!dstLastPrefixInt.valid) return V3(0.0f);
    const Inter srcLastPrefixInt = rehydrate(s, src.lastPrefixInt);
    if (!srcLastPrefixInt.valid) return V3(0.0f);
    Inter srcReconInt = rehydrate(s, src.reconInt);
    if (!srcReconInt.valid) return V3(0.0f);
    const uint32_t reconVertCount = (uint32_t)std::max<int>((int)(src.reconIdx() - 1), 0);
    Inter currInt = dstLastPrefixInt;
    PCG32 srcRng = src.rng;
    V3 throughput(1.0f);
    for (uint32_t idx = 0; idx < reconVertCount; ++idx) {
        srcRng.next_float();   // :43 dummy draw
        const BrdfSample bs = ggx_sample(s, currInt, srcRng);
        if (bs.samplingPdf <= 0.0f) return V3(0.0f);
        throughput *= bs.brdfVal;
        const V3 reconDir = bs.outDir;
        const V3 oldPos = currInt.pos;
        trace_surface(s, oldPos, reconDir, 1e-3f, 1e16f, currInt);
        if (!currInt.valid) return V3(0.0f);
    }
    const V3 reconVec = srcReconInt.pos - currInt.pos;
    const float reconLen = length(reconVec);
    const V3 reconDir = normalize(reconVec);
    if (any_nan_inf(reconDir)) return V3(0.0f);
    if (trace_occ(s, currInt.pos, reconDir, 1e-3f, reconLen)) return V3(0.0f);
    const BrdfEval b1 = ggx_eval(s, currInt, reconDir);
    if (b1.samplingPdf <= 0.0f) return V3(0.0f);
    throughput *= b1.brdfResult;
    if (src.reconIdx() < src.length()) {
        srcReconInt.inDir = reconDir;
        if (src.reconOutDir == V3(0.0f)) return V3(0.0f);
        const BrdfEval b2 = ggx_eval(s, srcReconInt, src.reconOutDir);
        if (b2.samplingPdf <= 0.0f) return V3(0.0f);
        throughput *= b2.brdfResult;
    }
    jacobian = reconnection_jacobian(srcLastPrefixInt.pos, currInt.pos, srcReconInt.pos, srcReconInt.normal);
    return throughput * src.postReconF;
}

static V3 shift_suffix(const Scene& s, const Suffix& src, const Suffix& dst, float& jacobian) {   // :111-123
    const Inter dstLast = rehydrate(s, dst.lastPrefixInt);
    if (!dstLast.valid) return V3(0.0f);
    return shift_suffix(s, src, dstLast, jacobian);
}

// ---------------------------------------------------------------------------------------
// Renderer state and passes
// ---------------------------------------------------------------------------------------

struct Renderer {
    const Scene* scene = nullptr;
    uint32_t W = 0, H = 0;
    Params P;
    std::vector<Reservoir<Prefix>> prefixRes;   // [2*P], interleaved like the reference
    std::vector<Prefix> canonPrefix;
    std::vector<Reservoir<Suffix>> suffixRes;
    std::vector<Suffix> canonSuffix;
    std::vector<GBuf> gbuf;
    std::vector<float> entryBox;                // [P][6] OptixAabb minX..maxZ (prefix_store_entries.cu)
    std::vector<uint8_t> entryValid;
    std::vector<Neighbor> neighbors;            // [P][k]
    std::vector<uint32_t> neighCount;
    PCG32 frameRng;
    uint64_t frameIdx = 0;
    uint32_t statMin = 0xFFFFFFFFu, statMax = 0;
    uint64_t statTotal = 0;
    double passMs[CX_PASS_COUNT] = {0};
    uint64_t passRays[2 * CX_PASS_COUNT] = {0};

    uint32_t pix() const { return W * H; }
};

static inline bool pixel_valid(uint32_t x, uint32_t y, const Renderer& R) { return x < R.W && y < R.H; }

static void write_output(Renderer& R, float* out, uint32_t p, const V3& rad) {
    V3 o = rad;
    if (R.P.lp.enableAccum) {
        const V3 old(out[p * 3 + 0], out[p * 3 + 1], out[p * 3 + 2]);
        const float blend = 1.0f / (float)(R.P.lp.frameIdx + 1);
        o = blend * rad + (1.0f - blend) * old;
    }
    out[p * 3 + 0] = o.x; out[p * 3 + 1] = o.y; out[p * 3 + 2] = o.z;
}

static void camera_ray(const Renderer& R, uint32_t x, uint32_t y, PCG32& rng, V3& origin, V3& dir) {
    // prefix_gen_temp_reuse.cu:122-126, include/graph/Camera.h:14-22
    const V2 j = rng.next2d();
    float ux = ((float)x + j.x) / (float)R.W;
    float uy = ((float)y + j.y) / (float)R.H;
    ux = 2.0f * ux - 1.0f;
    uy = 2.0f * uy - 1.0f;
    origin = R.P.camPos;
    dir = normalize(R.P.camU * ux + R.P.camV * uy + R.P.camW);
}

// PrefixGenTempReuse (shader/prefix_gen_temp_reuse.cu:11-98)
static bool prefix_gen_temp_reuse(Renderer& R, uint32_t p, const V3& origin, const V3& dir, PCG32& rng) {
    const Scene& s = *R.scene;
    const Params& P = R.P;
    trace_prefix(s, R.canonPrefix[p], origin, dir, rng, P);
    const Prefix& canon = R.canonPrefix[p];
    if (!canon.valid()) return false;
    const float canonPHat = luminance(canon.f);
    Reservoir<Prefix>& curr = R.prefixRes[2 * p + P.front];
    curr.reset();
    if (!P.lp.prefixEnableTemporal) {   // prevPixelCoord == pixelCoord is always valid
        curr.update(canon, canonPHat / canon.p, rng, 1.0f);
        curr.finalize();
        return true;
    }
    const Reservoir<Prefix>& prevRes = R.prefixRes[2 * p + P.back];
    const Prefix& prev = prevRes.sample;
    const bool skipBecauseOfNee = is_cond(P) && prev.nee();
    if (prevRes.wSum <= 0.0f || !prev.valid() || skipBecauseOfNee) {
        curr.update(canon, canonPHat / canon.p, rng, 1.0f);
        curr.finalize();
        return true;
    }
    float jCanonToPrev = 0.0f, jPrevToCanon = 0.0f;
    const V3 fFromCanonOfPrev = shift_prefix(s, prev, canon, jPrevToCanon);
    const V3 fFromPrevOfCanon = shift_prefix(s, canon, prev, jCanonToPrev);
    const float pFromCanonOfPrev = luminance(fFromCanonOfPrev) * jPrevToCanon;
    const float pairwiseK = 1.0f;
    const float prevMisWeight = neighbor_pairwise_mis(fFromCanonOfPrev, prev.f, jPrevToCanon, pairwiseK, 1.0f,
                                                      prevRes.confidence);
    const float canonMisWeight = 1.0f + canonical_pairwise_mis(canon.f, fFromPrevOfCanon, jCanonToPrev, pairwiseK,
                                                                1.0f, prevRes.confidence);
    const float canonRisWeight = canonMisWeight * canonPHat / canon.p;
    curr.update(canon, canonRisWeight, rng, 1.0f);
    const float prevRisWeight = prevMisWeight * pFromCanonOfPrev * prevRes.wSum;
    const Prefix shifted = prefix_shifted(s, prev, fFromCanonOfPrev, canon.primaryInt);
    curr.update(shifted, prevRisWeight, rng, prevRes.confidence);
    if (curr.wSum > 0.0f) {
        curr.wSum /= pairwiseK + 1.0f;
        curr.finalize();
    }
    return true;
}

// __raygen__prefix_gen_temp_reuse (shader/prefix_gen_temp_reuse.cu:100-193)
static void pass_prefix_gen_temp(Renderer& R, float* out) {
    const Params& P = R.P;
    const int64_t n = R.pix();
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t pi = 0; pi < n; pi++) {
        const uint32_t p = (uint32_t)pi, x = p % R.W, y = p / R.W;
        PCG32 rng(sample_tea64(p, P.lp.random));
        V3 origin, dir;
        camera_ray(R, x, y, rng, origin, dir);
        if (P.lp.rendererType == CX_CONDITIONAL_RESTIR || P.lp.rendererType == CX_RESTIR_PT) {
            const bool validPrimary = prefix_gen_temp_reuse(R, p, origin, dir, rng);
            if (is_restir_pt(P) && !P.lp.prefixEnableSpatial) {
                V3 rad(0.0f);
                if (validPrimary) {
                    const Reservoir<Prefix>& pr = R.prefixRes[2 * p + P.front];
                    rad = pr.wSum * pr.sample.f;
                }
                write_output(R, out, p, rad);
            }
            if ((is_restir_pt(P) && P.lp.prefixEnableSpatial) || (is_cond(P) && P.lp.suffixEnableSpatial))
                R.prefixRes[2 * p + P.back] = R.prefixRes[2 * p + P.front];
            GBuf g;
            g.prevX = x; g.prevY = y; g.rng = rng; g.primaryIntValid = validPrimary;
            R.gbuf[p] = g;
        } else if (P.lp.rendererType == CX_PATH_TRACER) {
            V3 rad = trace_complete_path(*R.scene, origin, dir, rng, P);
            if (any_nan_inf(rad)) rad = V3(0.0f);
            write_output(R, out, p, rad);
        }
    }
}

// PrefixSpatialReuse (shader/prefix_spatial_reuse.cu:11-134), reading neighbours from
// the snapshot `snap` of the back buffer taken at pass entry (C1).
static bool prefix_spatial_reuse(Renderer& R, uint32_t p, const std::vector<Reservoir<Prefix>>& snap) {
    const Scene& s = *R.scene;
    const Params& P = R.P;
    const uint32_t px = p % R.W, py = p / R.W;
    Reservoir<Prefix>& curr = R.prefixRes[2 * p + P.front];
    const Prefix currPrefix = curr.sample;
    if (!currPrefix.valid()) return false;
    const Prefix& canon = R.canonPrefix[p];
    if (!canon.valid() || !canon.primaryInt.valid) return false;
    PCG32& rng = R.gbuf[p].rng;
    const float currResWSum = curr.wSum;
    const float currResConfidence = curr.confidence;
    curr.reset();
    const int neighCount = P.lp.prefixSpatialCount;
    float canonMisWeight = 1.0f;
    uint32_t validNeighCount = 0, validNeighFlags = 0;
    uint32_t nxs[16], nys[16];
    for (uint32_t i = 0; i < (uint32_t)neighCount; ++i) {
        select_spatial_neighbor(px, py, P.lp.windowRadius, rng, nxs[i], nys[i]);
        if (!pixel_valid(nxs[i], nys[i], R)) continue;
        const uint32_t np = nys[i] * R.W + nxs[i];
        if (!R.gbuf[np].primaryIntValid) continue;
        const Prefix& np_ = snap[np].sample;
        const bool skip = is_cond(P) && np_.nee();
        if (!np_.valid() || skip) continue;
        ++validNeighCount;
        validNeighFlags |= 1u << i;
    }
    const float pairwiseK = (float)validNeighCount;
    for (uint32_t i = 0; i < (uint32_t)neighCount; ++i) {
        if (!((validNeighFlags >> i) & 1u)) continue;
        const uint32_t np = nys[i] * R.W + nxs[i];
        const Reservoir<Prefix>& nres = snap[np];
        const Prefix& neigh = nres.sample;
        float jNeighToCanon = 0.0f;
        const V3 fFromCanonOfNeigh = shift_prefix(s, neigh, currPrefix, jNeighToCanon);
        const float pFromCanonOfNeigh = luminance(fFromCanonOfNeigh) * jNeighToCanon;
        float jCanonToNeigh = 0.0f;
        const V3 fFromNeighOfCanon = shift_prefix(s, currPrefix, neigh, jCanonToNeigh);
        float neighMisWeight = neighbor_pairwise_mis(fFromCanonOfNeigh, neigh.f, jNeighToCanon, pairwiseK,
                                                     currResConfidence, nres.confidence);
        if (std::isnan(neighMisWeight) || std::isinf(neighMisWeight)) neighMisWeight = 0.0f;
        const float neighRisWeight = neighMisWeight * pFromCanonOfNeigh * nres.wSum;
        curr.update(prefix_shifted(s, neigh, fFromCanonOfNeigh, currPrefix.primaryInt), neighRisWeight, rng,
                    nres.confidence);
        canonMisWeight += canonical_pairwise_mis(currPrefix.f, fFromNeighOfCanon, jCanonToNeigh, pairwiseK,
                                                 currResConfidence, nres.confidence);
    }
    if (std::isinf(canonMisWeight) || std::isnan(canonMisWeight)) canonMisWeight = 0.0f;
    const float canonRisWeight = canonMisWeight * currResWSum * luminance(currPrefix.f);
    curr.update(currPrefix, canonRisWeight, rng, currResConfidence);
    if (curr.wSum > 0.0f) {
        curr.wSum /= pairwiseK + 1.0f;
        curr.finalize();
    }
    return true;
}

static void pass_prefix_spatial(Renderer& R, float* out) {   // prefix_spatial_reuse.cu:136-177
    const Params& P = R.P;
    const int64_t n = R.pix();
    std::vector<Reservoir<Prefix>> snap(n);
    for (int64_t p = 0; p < n; p++) snap[p] = R.prefixRes[2 * p + P.back];
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t pi = 0; pi < n; pi++) {
        const uint32_t p = (uint32_t)pi;
        V3 rad(0.0f);
        if (prefix_spatial_reuse(R, p, snap) && is_restir_pt(P)) {
            const Reservoir<Prefix>& pr = R.prefixRes[2 * p + P.front];
            rad = pr.wSum * pr.sample.f;
        }
        if (!is_restir_pt(P) || P.lp.spatialRoundIdx != P.lp.prefixSpatialRounds - 1) continue;
        write_output(R, out, p, rad);
    }
    for (int64_t p = 0; p < n; p++) R.prefixRes[2 * p + P.back] = R.prefixRes[2 * p + P.front];
}

static void invalidate_suffix(Renderer& R, uint32_t p) {   // suffix_gen_temp_reuse.cu:12-17
    Reservoir<Suffix>& r = R.suffixRes[2 * p + R.P.front];
    r.reset();
    r.sample.setValid(false);
}

// SuffixGenTempReuse (shader/suffix_gen_temp_reuse.cu:19-112)
static void suffix_gen_temp_reuse(Renderer& R, uint32_t p) {
    const Scene& s = *R.scene;
    const Params& P = R.P;
    const Prefix& prefix = R.prefixRes[2 * p + P.front].sample;
    if (!prefix.valid() || prefix.nee()) { invalidate_suffix(R, p); return; }
    PCG32& rng = R.gbuf[p].rng;
    trace_suffix(s, R.canonSuffix[p], prefix, rng, P);
    const Suffix& canon = R.canonSuffix[p];
    if (!canon.valid()) { invalidate_suffix(R, p); return; }
    const float canonPHat = luminance(canon.f);
    Reservoir<Suffix>& curr = R.suffixRes[2 * p + P.front];
    curr.reset();
    const uint32_t prevX = R.gbuf[p].prevX, prevY = R.gbuf[p].prevY;
    const uint32_t prevP = prevY * R.W + prevX;
    if (!P.lp.suffixEnableTemporal || !pixel_valid(prevX, prevY, R)) {
        curr.update(canon, canonPHat / canon.p, rng, 1.0f);
        curr.finalize();
        return;
    }
    const Reservoir<Suffix>& prevRes = R.suffixRes[2 * prevP + P.back];
    const Suffix& prev = prevRes.sample;
    if (prevRes.wSum <= 0.0f || !prev.valid()) {
        curr.update(canon, canonPHat / canon.p, rng, 1.0f);
        curr.finalize();
        return;
    }
    float jCanonToPrev = 0.0f, jPrevToCanon = 0.0f;
    const V3 fFromCanonOfPrev = shift_suffix(s, prev, canon, jPrevToCanon);
    const V3 fFromPrevOfCanon = shift_suffix(s, canon, prev, jCanonToPrev);
    const float pFromCanonOfPrev = luminance(fFromCanonOfPrev) * jPrevToCanon;
    const float pairwiseK = 1.0f;
    const float prevConfidence = prevRes.confidence;
    const float prevMisWeight = neighbor_pairwise_mis(fFromCanonOfPrev, prev.f, jPrevToCanon, pairwiseK, 1.0f,
                                                      prevConfidence);
    const float canonMisWeight = 1.0f + canonical_pairwise_mis(canon.f, fFromPrevOfCanon, jCanonToPrev, pairwiseK,
                                                                1.0f, prevConfidence);
    const float canonRisWeight = canonMisWeight * canonPHat / canon.p;
    curr.update(canon, canonRisWeight, rng, 1.0f);
    const float prevRisWeight = prevMisWeight * pFromCanonOfPrev * prevRes.wSum;
    curr.update(suffix_shifted(prev, canon.lastPrefixInt, fFromCanonOfPrev), prevRisWeight, rng, prevConfidence);
    if (curr.wSum > 0.0f) {
        curr.wSum /= pairwiseK + 1.0f;
        curr.finalize();
    }
}

static void pass_suffix_gen_temp(Renderer& R) {   // suffix_gen_temp_reuse.cu:114-134
    const Params& P = R.P;
    const int64_t n = R.pix();
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t pi = 0; pi < n; pi++) {
        const uint32_t p = (uint32_t)pi;
        suffix_gen_temp_reuse(R, p);
        R.suffixRes[2 * p + P.back] = R.suffixRes[2 * p + P.front];
    }
}

// SuffixSpatialReuse (shader/suffix_spatial_reuse.cu:16-145); both neighbour loops read the
// pass-entry snapshot (C1), which equals the front buffer at entry.
static void suffix_spatial_reuse(Renderer& R, uint32_t p, const std::vector<Reservoir<Suffix>>& snap) {
    const Scene& s = *R.scene;
    const Params& P = R.P;
    const uint32_t px = p % R.W, py = p / R.W;
    const Prefix& prefix = R.prefixRes[2 * p + P.front].sample;
    if (!prefix.valid() || prefix.nee()) { invalidate_suffix(R, p); return; }
    Reservoir<Suffix>& curr = R.suffixRes[2 * p + P.front];
    const Suffix currSuffix = curr.sample;
    if (!currSuffix.valid()) { invalidate_suffix(R, p); return; }
    const Suffix& canon = R.canonSuffix[p];
    if (!canon.valid()) { invalidate_suffix(R, p); return; }
    PCG32& rng = R.gbuf[p].rng;
    const float currResWSum = curr.wSum;
    const float currResConfidence = curr.confidence;
    curr.reset();
    const int neighCount = P.lp.suffixSpatialCount;
    float canonMisWeight = 1.0f;
    uint32_t validNeighCount = 0, validNeighFlags = 0;
    uint32_t nxs[16], nys[16];
    for (uint32_t i = 0; i < (uint32_t)neighCount; ++i) {
        select_spatial_neighbor(px, py, P.lp.windowRadius, rng, nxs[i], nys[i]);
        if (!pixel_valid(nxs[i], nys[i], R)) continue;
        const uint32_t np = nys[i] * R.W + nxs[i];
        if (!R.gbuf[np].primaryIntValid) continue;
        if (!snap[np].sample.valid()) continue;
        ++validNeighCount;
        validNeighFlags |= 1u << i;
    }
    const float pairwiseK = (float)validNeighCount;
    for (uint32_t i = 0; i < (uint32_t)neighCount; ++i) {
        if (!((validNeighFlags >> i) & 1u)) continue;
        const uint32_t np = nys[i] * R.W + nxs[i];
        const Reservoir<Suffix>& nres = snap[np];
        const Suffix& neigh = nres.sample;
        float jNeighToCanon = 0.0f;
        const V3 fFromCanonOfNeigh = shift_suffix(s, neigh, currSuffix, jNeighToCanon);
        const float pFromCanonOfNeigh = luminance(fFromCanonOfNeigh) * jNeighToCanon;
        float jCanonToNeigh = 0.0f;
        const V3 fFromNeighOfCanon = shift_suffix(s, currSuffix, neigh, jCanonToNeigh);
        float neighMisWeight = neighbor_pairwise_mis(fFromCanonOfNeigh, neigh.f, jNeighToCanon, pairwiseK,
                                                     currResConfidence, nres.confidence);
        if (std::isnan(neighMisWeight) || std::isinf(neighMisWeight)) neighMisWeight = 0.0f;
        const float neighRisWeight = neighMisWeight * pFromCanonOfNeigh * nres.wSum;
        curr.update(suffix_shifted(neigh, currSuffix.lastPrefixInt, fFromCanonOfNeigh), neighRisWeight, rng,
                    nres.confidence);
        canonMisWeight += canonical_pairwise_mis(currSuffix.f, fFromNeighOfCanon, jCanonToNeigh, pairwiseK,
                                                 currResConfidence, nres.confidence);
    }
    if (std::isinf(canonMisWeight) || std::isnan(canonMisWeight)) canonMisWeight = 0.0f;
    const float canonRisWeight = canonMisWeight * currResWSum * luminance(currSuffix.f);
    curr.update(currSuffix, canonRisWeight, rng, currResConfidence);
    if (curr.wSum > 0.0f) {
        curr.wSum /= pairwiseK + 1.0f;
        curr.finalize();
    }
}

static void pass_suffix_spatial(Renderer& R) {   // suffix_spatial_reuse.cu:147-167
    const Params& P = R.P;
    const int64_t n = R.pix();
    std::vector<Reservoir<Suffix>> snap(n);
    for (int64_t p = 0; p < n; p++) snap[p] = R.suffixRes[2 * p + P.back];
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t pi = 0; pi < n; pi++) suffix_spatial_reuse(R, (uint32_t)pi, snap);
    for (int64_t p = 0; p < n; p++) R.suffixRes[2 * p + P.back] = R.suffixRes[2 * p + P.front];
}

static void pass_store_entries(Renderer& R) {   // shader/prefix_store_entries.cu:9-53
    const Params& P = R.P;
    const int64_t n = R.pix();
#pragma omp parallel for schedule(static)
    for (int64_t pi = 0; pi < n; pi++) {
        const uint32_t p = (uint32_t)pi;
        const Prefix& prefix = R.prefixRes[2 * p + P.front].sample;
        const Suffix& suffix = R.suffixRes[2 * p + P.front].sample;
        const Inter last = rehydrate(*R.scene, prefix.lastInt);
        float* b = &R.entryBox[p * 6];
        if (prefix.valid() && suffix.valid() && last.valid) {
            float radius = P.lp.gatherRadius;
            if (P.lp.gatherRadiusType == CX_RADIUS_PATH_LENGTH) radius *= prefix.pathLen / 4.5f;
            const V3& pos = last.pos;
            b[0] = pos.x - radius; b[1] = pos.y - radius; b[2] = pos.z - radius;
            b[3] = pos.x + radius; b[4] = pos.y + radius; b[5] = pos.z + radius;
            R.entryValid[p] = 1;
        } else {
            for (int k = 0; k < 6; k++) b[k] = 0.0f;
            R.entryValid[p] = 0;
        }
    }
}

// ---------------------------------------------------------------------------------------
// Neighbour search (shader/prefix_search_hit.cu:9-73 under C2)
// ---------------------------------------------------------------------------------------

struct EntrySet {
    std::vector<V3> pos;
    std::vector<float> radius;    // (maxX - minX) / 2, prefix_search_hit.cu:33-34
    std::vector<uint8_t> valid;
    // uniform grid over valid entries
    V3 gmin;
    float cell = 1.0f;
    int nx = 0, ny = 0, nz = 0;
    std::vector<uint32_t> cellStart, cellItems;
    float rmax = 0.0f;
};

static void entries_build(EntrySet& E) {
    const uint32_t n = (uint32_t)E.pos.size();
    V3 lo(INFINITY), hi(-INFINITY);
    E.rmax = 0.0f;
    uint32_t nv = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (!E.valid[i]) continue;
        nv++;
        lo = V3(std::min(lo.x, E.pos[i].x), std::min(lo.y, E.pos[i].y), std::min(lo.z, E.pos[i].z));
        hi = V3(std::max(hi.x, E.pos[i].x), std::max(hi.y, E.pos[i].y), std::max(hi.z, E.pos[i].z));
        E.rmax = std::max(E.rmax, E.radius[i]);
    }
    E.cellStart.clear();
    E.cellItems.clear();
    if (nv == 0) { E.nx = E.ny = E.nz = 0; return; }
    const float ext = std::max(std::max(hi.x - lo.x, hi.y - lo.y), hi.z - lo.z);
    float cell = std::max(E.rmax, ext / 256.0f);
    if (!(cell > 0.0f)) cell = 1.0f;
    E.cell = cell;
    E.gmin = lo;
    E.nx = std::min(1024, (int)((hi.x - lo.x) / cell) + 1);
    E.ny = std::min(1024, (int)((hi.y - lo.y) / cell) + 1);
    E.nz = std::min(1024, (int)((hi.z - lo.z) / cell) + 1);
    const size_t nc = (size_t)E.nx * E.ny * E.nz;
    std::vector<uint32_t> cnt(nc + 1, 0);
    auto cell_of = [&](const V3& q, int& cx, int& cy, int& cz) {
        cx = std::min(E.nx - 1, std::max(0, (int)((q.x - lo.x) / cell)));
        cy = std::min(E.ny - 1, std::max(0, (int)((q.y - lo.y) / cell)));
        cz = std::min(E.nz - 1, std::max(0, (int)((q.z - lo.z) / cell)));
    };
    std::vector<uint32_t> cid(n, 0);
    for (uint32_t i = 0; i < n; i++) {
        if (!E.valid[i]) continue;
        int cx, cy, cz;
        cell_of(E.pos[i], cx, cy, cz);
        cid[i] = (uint32_t)(((size_t)cz * E.ny + cy) * E.nx + cx);
        cnt[cid[i] + 1]++;
    }
    for (size_t c = 0; c < nc; c++) cnt[c + 1] += cnt[c];
    E.cellStart = cnt;
    E.cellItems.assign(nv, 0);
    std::vector<uint32_t> fill(cnt.begin(), cnt.end() - 1);
    for (uint32_t i = 0; i < n; i++)
        if (E.valid[i]) E.cellItems[fill[cid[i]]++] = i;
}

// Candidate test of prefix_search_hit.cu:18-42 (validity folded into E.valid) and the
// (distance, index) ordering of C2.
static uint32_t entries_query(const EntrySet& E, const V3& q, uint32_t k, Neighbor* out) {
    if (k == 0 || E.nx == 0) return 0;
    std::vector<Neighbor> best;
    best.reserve(k + 1);
    auto push = [&](uint32_t i, float d) {
        auto less = [](const Neighbor& a, const Neighbor& b) {
            return a.distance < b.distance || (a.distance == b.distance && a.pixelIdx < b.pixelIdx);
        };
        Neighbor nb{i, d};
        if (best.size() == k && !less(nb, best.back())) return;
        auto it = std::upper_bound(best.begin(), best.end(), nb, less);
        best.insert(it, nb);
        if (best.size() > k) best.pop_back();
    };
    if (!(q.x == q.x) || !(q.y == q.y) || !(q.z == q.z)) return 0;
    const float r = E.rmax;
    const int x0 = std::max(0, (int)std::floor((q.x - r - E.gmin.x) / E.cell) - 1);
    const int x1 = std::min(E.nx - 1, (int)std::floor((q.x + r - E.gmin.x) / E.cell) + 1);
    const int y0 = std::max(0, (int)std::floor((q.y - r - E.gmin.y) / E.cell) - 1);
    const int y1 = std::min(E.ny - 1, (int)std::floor((q.y + r - E.gmin.y) / E.cell) + 1);
    const int z0 = std::max(0, (int)std::floor((q.z - r - E.gmin.z) / E.cell) - 1);
    const int z1 = std::min(E.nz - 1, (int)std::floor((q.z + r - E.gmin.z) / E.cell) + 1);
    for (int cz = z0; cz <= z1; cz++)
        for (int cy = y0; cy <= y1; cy++)
            for (int cx = x0; cx <= x1; cx++) {
                const size_t c = ((size_t)cz * E.ny + cy) * E.nx + cx;
                for (uint32_t j = E.cellStart[c]; j < E.cellStart[c + 1]; j++) {
                    const uint32_t i = E.cellItems[j];
                    const float d = distance(q, E.pos[i]);
                    if (d > E.radius[i]) continue;
                    push(i, d);
                }
            }
    for (size_t i = 0; i < best.size(); i++) out[i] = best[i];
    return (uint32_t)best.size();
}

static uint32_t entries_query_brute(const EntrySet& E, const V3& q, uint32_t k, Neighbor* out) {
    std::vector<Neighbor> all;
    for (uint32_t i = 0; i < (uint32_t)E.pos.size(); i++) {
        if (!E.valid[i]) continue;
        const float d = distance(q, E.pos[i]);
        if (d > E.radius[i]) continue;
        all.push_back({i, d});
    }
    std::sort(all.begin(), all.end(), [](const Neighbor& a, const Neighbor& b) {
        return a.distance < b.distance || (a.distance == b.distance && a.pixelIdx < b.pixelIdx);
    });
    const uint32_t m = (uint32_t)std::min<size_t>(k, all.size());
    for (uint32_t i = 0; i < m; i++) out[i] = all[i];
    return m;
}

static void renderer_entries(Renderer& R, EntrySet& E) {
    const uint32_t n = R.pix();
    E.pos.resize(n);
    E.radius.resize(n);
    E.valid.resize(n);
    for (uint32_t p = 0; p < n; p++) {
        const float* b = &R.entryBox[p * 6];
        E.valid[p] = R.entryValid[p];
        E.radius[p] = (b[3] - b[0]) / 2.0f;
        if (E.valid[p]) E.pos[p] = rehydrate(*R.scene, R.prefixRes[2 * p + R.P.front].sample.lastInt).pos;
    }
    entries_build(E);
}

// ---------------------------------------------------------------------------------------
// Final gather (shader/final_gather.cu:69-356)
// ---------------------------------------------------------------------------------------

// One valid neighbour of the canonical gather: its suffix shifted onto the canonical prefix
// (value, Jacobian), its own value, reservoir weight and confidence, and the canonical suffix
// shifted onto its prefix (value, Jacobian).
struct GatherTerm {
    V3 fFromCanonOfNeigh;
    float jNeighToCanon = 0.0f;
    V3 neighF;
    float neighWSum = 0.0f, neighConf = 0.0f;
    V3 fFromNeighOfCanon;
    float jCanonToNeigh = 0.0f;
};

// The pairwise-MIS combination of final_gather.cu:135-187 over `n` valid neighbours.
static V3 gather_combine(const V3& canonPreF, float canonPreP, const V3& canonSufF, const V3& canonContrib,
                         const GatherTerm* t, uint32_t n) {
    V3 suffixContrib(0.0f);
    const float pairwiseK = (float)n;
    float canonMisWeight = 1.0f;
    for (uint32_t i = 0; i < n; ++i) {
        float neighMisWeight = neighbor_pairwise_mis(t[i].fFromCanonOfNeigh, t[i].neighF, t[i].jNeighToCanon,
                                                     pairwiseK, 1.0f, t[i].neighConf);
        if (std::isnan(neighMisWeight) || std::isinf(neighMisWeight)) neighMisWeight = 0.0f;
        const V3 neighContrib = neighMisWeight * t[i].fFromCanonOfNeigh * t[i].neighWSum * t[i].jNeighToCanon;
        if (!any_nan_inf(neighContrib)) suffixContrib += neighContrib;
        canonMisWeight += canonical_pairwise_mis(canonSufF, t[i].fFromNeighOfCanon, t[i].jCanonToNeigh, pairwiseK,
                                                 1.0f, t[i].neighConf);
    }
    V3 gatherContrib = suffixContrib * canonPreF / canonPreP;
    gatherContrib /= (float)(n + 1);
    V3 result = gatherContrib + (canonContrib * canonMisWeight);
    result /= pairwiseK + 1.0f;
    if (any_nan_inf(result)) return V3(0.0f);
    return result;
}

static V3 final_gather_canon(Renderer& R, const EntrySet& E, uint32_t p, const V3& origin, const V3& dir,
                             uint32_t k, PCG32& rng) {   // final_gather.cu:69-188
    const Scene& s = *R.scene;
    const Params& P = R.P;
    Prefix canonPrefix;
    trace_prefix(s, canonPrefix, origin, dir, rng, P);
    if (!canonPrefix.valid()) return V3(0.0f);
    if (canonPrefix.nee()) return canonPrefix.f / canonPrefix.p;
    Suffix canonSuffix;
    trace_suffix(s, canonSuffix, canonPrefix, rng, P);
    if (!canonSuffix.valid()) return V3(0.0f);
    const V3 canonContrib = canonPrefix.f * canonSuffix.f / (canonPrefix.p * canonSuffix.p);
    if (k == 0) return canonContrib;
    const Inter lastPrefixInt = rehydrate(s, canonSuffix.lastPrefixInt);
    Neighbor* nb = &R.neighbors[(size_t)p * k];
    const uint32_t neighCount = entries_query(E, lastPrefixInt.pos, k, nb);
    R.neighCount[p] = neighCount;
    uint32_t validNeighCount = 0, validNeighFlags = 0;
    for (uint32_t i = 0; i < neighCount; ++i) {
        const Suffix& ns = R.suffixRes[2 * nb[i].pixelIdx + P.front].sample;
        if (!ns.valid()) continue;
        ++validNeighCount;
        validNeighFlags |= 1u << i;
    }
    GatherTerm terms[32];
    uint32_t n = 0;
    for (uint32_t i = 0; i < neighCount; ++i) {
        if (!(validNeighFlags & (1u << i))) continue;
        const Reservoir<Suffix>& nres = R.suffixRes[2 * nb[i].pixelIdx + P.front];
        GatherTerm& t = terms[n++];
        t.fFromCanonOfNeigh = shift_suffix(s, nres.sample, canonSuffix, t.jNeighToCanon);
        t.fFromNeighOfCanon = shift_suffix(s, canonSuffix, nres.sample, t.jCanonToNeigh);
        t.neighF = nres.sample.f;
        t.neighWSum = nres.wSum;
        t.neighConf = nres.confidence;
    }
    return gather_combine(canonPrefix.f, canonPrefix.p, canonSuffix.f, canonContrib, terms, n);
}

static V3 final_gather_not_canon(Renderer& R, const EntrySet& E, uint32_t p, const V3& origin, const V3& dir,
                                 uint32_t k, PCG32& rng) {   // final_gather.cu:190-288
    const Scene& s = *R.scene;
    const Params& P = R.P;
    if (k == 0) return V3(0.0f);
    Inter lastPrefixInt;
    V3 prefixThroughput(0.0f);
    float prefixP = 0.0f;
    if (!trace_prefix_for_final_gather(s, prefixThroughput, prefixP, lastPrefixInt, origin, dir,
                                       (uint64_t)P.lp.prefixLen, rng, P))
        return V3(0.0f);
    Neighbor* nb = &R.neighbors[(size_t)p * k];
    const uint32_t neighCount = entries_query(E, lastPrefixInt.pos, k, nb);
    R.neighCount[p] = neighCount;
    uint32_t validNeighCount = 0, validNeighFlags = 0;
    for (uint32_t i = 0; i < neighCount; ++i) {
        if (!R.suffixRes[2 * nb[i].pixelIdx + P.front].sample.valid()) continue;
        ++validNeighCount;
        validNeighFlags |= 1u << i;
    }
    V3 suffixContrib(0.0f);
    const float pairwiseK = (float)validNeighCount;
    for (uint32_t i = 0; i < neighCount; ++i) {
        if (!(validNeighFlags & (1u << i))) continue;
        const Reservoir<Suffix>& nres = R.suffixRes[2 * nb[i].pixelIdx + P.front];
        const Suffix& neigh = nres.sample;
        float jNeighToCanon = 0.0f;
        const V3 fFromCanonOfNeigh = shift_suffix(s, neigh, lastPrefixInt, jNeighToCanon);
        float neighMisWeight = neighbor_pairwise_mis(fFromCanonOfNeigh, neigh.f, jNeighToCanon, pairwiseK, 1.0f,
                                                     nres.confidence);
        if (std::isnan(neighMisWeight) || std::isinf(neighMisWeight)) neighMisWeight = 0.0f;
        const V3 neighContrib = neighMisWeight * fFromCanonOfNeigh * nres.wSum * jNeighToCanon;
        if (!any_nan_inf(neighContrib)) suffixContrib += neighContrib;
    }
    V3 result = suffixContrib * prefixThroughput / prefixP;
    result /= (float)P.lp.gatherN;
    result /= pairwiseK + 1.0f;
    if (any_nan_inf(result)) return V3(0.0f);
    return result;
}

static void pass_final_gather(Renderer& R, const EntrySet& E, float* out) {   // final_gather.cu:290-356
    const Params& P = R.P;
    const int64_t n = R.pix();
    const uint32_t k = (uint32_t)(P.lp.gatherM - 1);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t pi = 0; pi < n; pi++) {
        const uint32_t p = (uint32_t)pi, x = p % R.W, y = p / R.W;
        V3 rad(0.0f);
        if (R.canonPrefix[p].primaryInt.valid) {
            PCG32& rng = R.gbuf[p].rng;
            V3 origin, dir;
            camera_ray(R, x, y, rng, origin, dir);
            rad += final_gather_canon(R, E, p, origin, dir, k, rng);
            for (int32_t g = 1; g < P.lp.gatherN; ++g) rad += final_gather_not_canon(R, E, p, origin, dir, k, rng);
            if (any_nan_inf(rad)) rad = V3(0.0f);
        }
        write_output(R, out, p, rad);
    }
    if (P.lp.trackPrefixStats) {
        // final_gather.cu:110-115 (only queries that ran are counted)
    }
}

static double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

static void end_pass(Renderer& R, int pass, std::chrono::steady_clock::time_point t0) {
    R.passMs[pass] += ms_since(t0);
    uint64_t c, s;
    take_rays(c, s);
    R.passRays[2 * pass] += c;
    R.passRays[2 * pass + 1] += s;
}

// Renderer::LaunchFrame (src/Renderer.cpp:209-406)
static void launch_frame(Renderer& R, const cx_camera& cam, bool camChanged, float* out) {
    Params& P = R.P;
    ++R.frameIdx;
    if (camChanged || !P.lp.enableAccum) R.frameIdx = 0;
    P.lp.frameIdx = R.frameIdx;
    P.lp.random = R.frameRng.next_u32();   // C4
    P.lp.width = R.W;
    P.lp.height = R.H;
    P.camPos = V3(cam.pos[0], cam.pos[1], cam.pos[2]);
    P.camU = V3(cam.U[0], cam.U[1], cam.U[2]);
    P.camV = V3(cam.V[0], cam.V[1], cam.V[2]);
    P.camW = V3(cam.W[0], cam.W[1], cam.W[2]);
    if (P.front == 0) { P.front = 1; P.back = 0; } else { P.front = 0; P.back = 1; }
    P.lp.spatialRoundIdx = 0;
    for (double& m : R.passMs) m = 0.0;
    for (uint64_t& c : R.passRays) c = 0;
    uint64_t dc, ds;
    take_rays(dc, ds);
    auto t0 = std::chrono::steady_clock::now();
    pass_prefix_gen_temp(R, out);
    end_pass(R, CX_PASS_PREFIX_GEN_TEMP, t0);
    if (P.lp.rendererType == CX_CONDITIONAL_RESTIR || P.lp.rendererType == CX_RESTIR_PT) {
        if (P.lp.prefixEnableSpatial) {
            t0 = std::chrono::steady_clock::now();
            for (int32_t r = 0; r < P.lp.prefixSpatialRounds; ++r) {
                P.lp.spatialRoundIdx = r;
                pass_prefix_spatial(R, out);
            }
            end_pass(R, CX_PASS_PREFIX_SPATIAL, t0);
        }
        if (is_cond(P)) {
            t0 = std::chrono::steady_clock::now();
            pass_suffix_gen_temp(R);
            end_pass(R, CX_PASS_SUFFIX_GEN_TEMP, t0);
            if (P.lp.suffixEnableSpatial && P.lp.suffixSpatialCount > 0) {
                t0 = std::chrono::steady_clock::now();
                for (int32_t r = 0; r < P.lp.suffixSpatialRounds; ++r) {
                    P.lp.spatialRoundIdx = r;
                    pass_suffix_spatial(R);
                }
                end_pass(R, CX_PASS_SUFFIX_SPATIAL, t0);
            }
            t0 = std::chrono::steady_clock::now();
            pass_store_entries(R);
            end_pass(R, CX_PASS_PREFIX_STORE_ENTRIES, t0);
            t0 = std::chrono::steady_clock::now();
            EntrySet E;
            renderer_entries(R, E);
            const uint32_t k = (uint32_t)std::max(0, P.lp.gatherM - 1);
            R.neighbors.assign((size_t)R.pix() * std::max(1u, k), Neighbor());
            R.neighCount.assign(R.pix(), 0);
            end_pass(R, CX_PASS_ENTRIES_BUILD, t0);
            t0 = std::chrono::steady_clock::now();
            pass_final_gather(R, E, out);
            end_pass(R, CX_PASS_FINAL_GATHER, t0);
        }
    }
}

static void renderer_reset(Renderer& R) {   // src/Renderer.cpp:155-199
    const uint32_t n = R.pix();
    R.prefixRes.assign(2 * (size_t)n, Reservoir<Prefix>());
    R.canonPrefix.assign(n, Prefix());
    R.suffixRes.assign(2 * (size_t)n, Reservoir<Suffix>());
    R.canonSuffix.assign(n, Suffix());
    R.gbuf.assign(n, GBuf());
    R.entryBox.assign((size_t)n * 6, 0.0f);
    R.entryValid.assign(n, 0);
    R.neighCount.assign(n, 0);
    R.P.front = 0;
    R.P.back = 0;
    R.frameIdx = 0;
}

}  // namespace orc

// =========================================================================================
// C API for ctypes (tests/, smoke, bench cpu_baseline)
// =========================================================================================

using namespace orc;

struct orc_packed_dump {   // flat view of Packed for tests
    uint32_t prim;
    float bu, bv;
    float inDir[3];
    uint32_t valid;
};

struct orc_prefix_dump {
    uint64_t rngState;
    uint32_t flags;
    float f[3], postReconF[3];
    float p, pathLen;
    float reconOutDir[3];
    float wSum, confidence;
    orc_packed_dump primaryInt, reconInt, lastInt;
};

struct orc_suffix_dump {
    uint64_t rngState;
    uint32_t flags;
    float f[3], postReconF[3];
    float p;
    float reconOutDir[3];
    float wSum, confidence;
    orc_packed_dump reconInt, lastPrefixInt;
};

static void dump_packed(const Packed& p, orc_packed_dump& d) {
    d.prim = p.prim; d.bu = p.bu; d.bv = p.bv;
    d.inDir[0] = p.inDir.x; d.inDir[1] = p.inDir.y; d.inDir[2] = p.inDir.z;
    d.valid = p.valid ? 1u : 0u;
}

static void v3out(const V3& v, float* o) { o[0] = v.x; o[1] = v.y; o[2] = v.z; }

// ---------------------------------------------------------------------------------------
// FP64 truth: the expectation of the PathTracer estimator (path_gen.h:389-472, the NEE model
// of include/graph/sample_emitter.h and include/model/Emitter.h:34-54), estimated with every
// shading quantity in double precision and independent 53-bit uniforms.  Only the traversal
// runs in float (geometry is float data); hit positions and normals are recomputed in double
// from the float vertices and the hit's barycentrics.
// ---------------------------------------------------------------------------------------

struct D3 {
    double x = 0, y = 0, z = 0;
    D3() = default;
    D3(double a, double b, double c) : x(a), y(b), z(c) {}
    explicit D3(const V3& v) : x(v.x), y(v.y), z(v.z) {}
    explicit D3(double s) : x(s), y(s), z(s) {}
    D3 operator+(const D3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    D3 operator-(const D3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    D3 operator*(const D3& o) const { return {x * o.x, y * o.y, z * o.z}; }
    D3 operator*(double s) const { return {x * s, y * s, z * s}; }
    D3 operator/(double s) const { return {x / s, y / s, z / s}; }
    D3 operator-() const { return {-x, -y, -z}; }
};
static inline D3 operator*(double s, const D3& v) { return v * s; }
static inline double ddot(const D3& a, const D3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double dlen(const D3& a) { return std::sqrt(ddot(a, a)); }
static inline D3 dnorm(const D3& a) { return a / dlen(a); }
static inline V3 tof(const D3& a) { return V3((float)a.x, (float)a.y, (float)a.z); }

// splitmix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number generators", 2014).
struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct DInter {
    bool valid = false;
    uint32_t prim = 0;
    float bu = 0.0f, bv = 0.0f;
    D3 pos, nrm, inDir;
};

static DInter d_trace(const Scene& s, const D3& o, const D3& d) {
    DInter it;
    it.inDir = d;
    const Hit h = trace_closest(s, tof(o), tof(d), 1e-3f, 1e16f);
    if (h.prim == INVALID_PRIM) return it;
    it.valid = true;
    it.prim = h.prim;
    it.bu = h.u;
    it.bv = h.v;
    const double u = h.u, v = h.v, w = 1.0 - u - v;
    const D3 p0(s.vtx(h.prim, 0)), p1(s.vtx(h.prim, 1)), p2(s.vtx(h.prim, 2));
    const D3 n0(s.nrm[s.idx[h.prim * 3]]), n1(s.nrm[s.idx[h.prim * 3 + 1]]), n2(s.nrm[s.idx[h.prim * 3 + 2]]);
    it.pos = w * p0 + u * p1 + v * p2;
    it.nrm = dnorm(w * n0 + u * n1 + v * n2);
    return it;
}

// A hit's material in double (brdf.cu:121-138): texture values are the filter's exact sums
// divided in double; diffParam / roughParam say whether the value is the material's constant.
struct DMat {
    D3 diff, spec, emissive;
    double rough = 0.0, metal = 0.0;
    bool diffParam = true, roughParam = true;
};

static D3 tex_fetch64(const Texture& t, float u, float v) {
    uint32_t q[3];
    texel_sum(t, u, v, q);
    return D3(q[0] / 16711680.0, q[1] / 16711680.0, q[2] / 16711680.0);
}

static DMat resolve64(const Scene& s, const Material& m, uint32_t prim, float bu, float bv) {
    DMat r;
    float u = 0.0f, v = 0.0f;
    if (m.diffTex >= 0 || m.specTex >= 0 || m.roughTex >= 0) inter_uv(s, prim, bu, bv, u, v);
    r.diff = m.diffTex >= 0 ? tex_fetch64(s.textures[m.diffTex], u, v) : D3(m.diff);
    r.diffParam = m.diffTex < 0;
    D3 spec = m.specTex >= 0 ? tex_fetch64(s.textures[m.specTex], u, v) : D3(m.spec);
    r.rough = m.roughTex >= 0 ? tex_fetch64(s.textures[m.roughTex], u, v).x : (double)m.roughness;
    r.roughParam = m.roughTex < 0;
    if (m.specUsage == 1) {
        r.rough = spec.y;
        r.metal = spec.z;
        spec = D3(1.0);
        r.roughParam = false;
    }
    r.spec = r.diff * (1.0 - r.metal) + spec * r.metal;
    r.emissive = D3(m.emissive);
    return r;
}

static DMat d_material(const Scene& s, const DInter& it) {
    return resolve64(s, s.mats[s.triMat[it.prim]], it.prim, it.bu, it.bv);
}

static inline double d_ggx_D(double nh, double r) {
    const double a2 = r * r;
    const double d = (nh * a2 - nh) * nh + 1.0;
    return a2 / (M_PI * d * d);
}
static inline double d_smith(double nl, double nv, double r) {
    const double a2 = r * r;
    return 0.5 / (nl * std::sqrt(nv * nv * (1 - a2) + a2) + nv * std::sqrt(nl * nl * (1 - a2) + a2));
}

// BSDF value (times cos) and sampling pdf of brdf.cu:115-190 for light direction l.
static void d_eval(const DMat& m, const DInter& it, const D3& l, D3& f, double& pdf) {
    const D3 v = -it.inDir;
    const double nv = ddot(it.nrm, v), nl = ddot(it.nrm, l);
    f = D3(0.0);
    pdf = 0.0;
    if (nv <= 0.0 || nl < 0.0) return;
    const double sd = m.diff.x + m.diff.y + m.diff.z, ss = m.spec.x + m.spec.y + m.spec.z;
    const double diffProb = sd / (sd + ss);
    const D3 h = dnorm(l + v);
    const double nh = ddot(h, it.nrm), lh = ddot(h, l);
    D3 spec(0.0);
    if (ddot(m.spec, m.spec) > 1e-6) {
        const double q = std::pow(std::max(0.0, 1.0 - lh), 5.0);
        const D3 F = m.spec + (D3(1.0) - m.spec) * q;
        spec = F * (d_ggx_D(nh, m.rough) * d_smith(nl, nv, m.rough));
    }
    f = (m.diff / M_PI + spec) * nl;
    pdf = diffProb * nl / M_PI + (1.0 - diffProb) * d_ggx_D(nh, m.rough) * nh / (4.0 * lh);
}

// Direction sampling of brdf.cu:214-338 (cosine lobe or GGX normal distribution).
static bool d_sample(const DMat& m, const DInter& it, SplitMix& rng, D3& l) {
    const D3 n = it.nrm;
    if (ddot(n, -it.inDir) <= 0.0) return false;
    const double sz = n.z >= 0 ? 1.0 : -1.0, a = 1.0 / (sz + n.z), ya = n.y * a, b = n.x * ya, c = n.x * sz;
    const D3 t0(c * n.x * a - 1, sz * b, c), t1(b, n.y * ya - sz, n.y);
    const double sd = m.diff.x + m.diff.y + m.diff.z, ss = m.spec.x + m.spec.y + m.spec.z;
    if (rng.uniform() < sd / (sd + ss)) {
        const double phi = 2.0 * M_PI * rng.uniform(), r = std::sqrt(rng.uniform());
        const double x = r * std::cos(phi), y = r * std::sin(phi), z = std::sqrt(std::max(0.0, 1.0 - x * x - y * y));
        l = t0 * x + t1 * y + n * z;
        return true;
    }
    const double u = rng.uniform();
    const double mz = std::sqrt(std::min(1.0, std::max(0.0, (1 - u) / (1 + (m.rough * m.rough - 1) * u))));
    const double mr = std::sqrt(std::max(0.0, 1 - mz * mz));
    const double phi = 2.0 * M_PI * rng.uniform();
    const D3 h = t0 * (mr * std::cos(phi)) + t1 * (mr * std::sin(phi)) + n * mz;
    l = it.inDir - h * (2.0 * ddot(h, it.inDir));
    return ddot(l, n) >= 0.0;
}

// One sample of the PathTracer estimator from the pixel's camera ray.
static D3 d_path(const Scene& s, const D3& origin, const D3& dir, uint32_t maxLen, double neeProb, SplitMix& rng) {
    D3 pos = origin, d = dir, thr(1.0);
    for (uint32_t depth = 1; depth <= maxLen; ++depth) {
        const DInter it = d_trace(s, pos, d);
        if (!it.valid) break;
        const DMat m = d_material(s, it);
        if (rng.uniform() < neeProb) {
            thr = thr / neeProb;
            const EmitterData& e = s.emitters[std::min<size_t>((size_t)(rng.uniform() * s.emitters.size()),
                                                               s.emitters.size() - 1)];
            const size_t nf = e.faces.size() / 3;
            const size_t fi = std::min<size_t>((size_t)(rng.uniform() * nf), nf - 1);
            double r0 = rng.uniform(), r1 = rng.uniform(), r2 = rng.uniform();
            const double rs = r0 + r1 + r2;
            r0 /= rs; r1 /= rs; r2 /= rs;
            const D3 lp = r0 * D3(e.faces[3 * fi]) + r1 * D3(e.faces[3 * fi + 1]) + r2 * D3(e.faces[3 * fi + 2]);
            const D3 lv = lp - it.pos;
            const double dist = dlen(lv);
            const D3 ldir = lv / dist;
            if (trace_occluded(s, tof(it.pos), tof(ldir), 1e-3f, (float)dist)) return D3(0.0);
            D3 f;
            double pdf;
            d_eval(m, it, ldir, f, pdf);
            D3 rad = thr * f * D3(e.color) / (dist * dist);
            if (depth == 1) rad = rad + m.emissive;
            return rad;
        }
        D3 l;
        if (!d_sample(m, it, rng, l)) break;
        D3 f;
        double pdf;
        d_eval(m, it, l, f, pdf);
        if (pdf <= 0.0) break;
        thr = thr * f / pdf / (1.0 - neeProb);
        pos = it.pos;
        d = l;
    }
    return D3(0.0);
}


// ---------------------------------------------------------------------------------------
// FP64 path-replay adjoint (checks cx_renderer_adjoint).  The estimator and the replay are
// those of crisoptix_amd/csrc/cx_adjoint.hip: the gradient of I = W * f(x) with the
// reservoir-selected path x and its weight W held fixed (ReSTIR PT), or of f(x)/p(x) with p
// held fixed (path tracer).  Geometry, directions and decisions are replayed in float with
// the sampling tables, exactly as the frame produced them; every BSDF value, emitter colour,
// product and derivative is computed in double from those float inputs and the evaluation
// tables.  The BSDF derivatives are written out from the formula of brdf.cu:99-190 with
// F0 = mix(albedo, specular, metal) and checked against double-precision central differences.
// ---------------------------------------------------------------------------------------

namespace {

struct AFactor {
    D3 v, dA, dR, dC;
    uint32_t mat = 0, emit = 0xFFFFFFFFu;
};

struct ADouble {
    D3 val, dA, dR;
};

ADouble brdf64(const DMat& m, const V3& nF, const V3& vF, const V3& lF) {
    const D3 a = m.diff, f0 = m.spec, n(nF), v(vF), l(lF);
    const double r = m.rough;
    const double nV = ddot(n, v), nL = ddot(n, l);
    const D3 h = dnorm(l + v);
    const double nH = ddot(h, n), lH = ddot(h, l);
    const double pi = 3.14159265358979323846;
    ADouble o;
    D3 spec(0.0), dR(0.0);
    double dAs = 1.0 / pi;
    if (ddot(f0, f0) > 1e-6) {   // F0 = albedo (1 - metal) + specular * metal
        const double a2 = r * r;
        const double q = nH * nH * (a2 - 1.0) + 1.0;
        const double D = a2 / (pi * q * q);
        const double sv = std::sqrt(nV * nV * (1.0 - a2) + a2), sl = std::sqrt(nL * nL * (1.0 - a2) + a2);
        const double lam = nL * sv + nV * sl;
        const double V = 0.5 / lam;
        const double x = std::max(0.0, 1.0 - lH);
        const double P = x * x * x * x * x;
        const D3 F = f0 + (D3(1.0) - f0) * P;
        spec = F * (D * V);
        dAs += D * V * (1.0 - P) * (1.0 - m.metal);
        const double dD = (q - 2.0 * a2 * nH * nH) / (pi * q * q * q);
        const double dlam = nL * (1.0 - nV * nV) / (2.0 * sv) + nV * (1.0 - nL * nL) / (2.0 * sl);
        const double dV = -0.5 * dlam / (lam * lam);
        dR = F * (2.0 * r * (dD * V + D * dV));
    }
    o.val = (a / pi + spec) * nL;
    o.dA = m.diffParam ? D3(dAs * nL) : D3(0.0);
    o.dR = m.roughParam ? dR * nL : D3(0.0);
    return o;
}

struct AdjOut {
    double* gA;
    double* gR;
    double* gE;
};

struct AReplay {
    const Scene& s;
    const Params& P;
    std::vector<AFactor> F;
    bool ok = true;
};

EmitterSample sample_emitter_idx(const Scene& s, PCG32& rng, uint32_t& ei) {
    const uint64_t count = s.emitters.size();
    const uint64_t e = rng.next_u64() % count;
    const EmitterData& ed = s.emitters[e];
    const uint64_t faceIdx = rng.next_u64() % (uint64_t)(ed.faces.size() / 3);
    const V3 v0 = ed.faces[faceIdx * 3 + 0], v1 = ed.faces[faceIdx * 3 + 1], v2 = ed.faces[faceIdx * 3 + 2];
    float r0 = rng.next_float();
    float r1 = rng.next_float();
    float r2 = rng.next_float();
    const float rSum = r0 + r1 + r2;
    r0 /= rSum;
    r1 /= rSum;
    r2 /= rSum;
    ei = (uint32_t)e;
    return {r0 * v0 + r1 * v1 + r2 * v2, ed.color, 1.0f};
}

// NEE at `it`: factor (brdf * colour) * distFactor in double; brdf/colour/distFactor/emission
// returned for the path tracer's order of terms.
bool nee64(AReplay& R, const Inter& it, PCG32& rng, bool requireValue, AFactor& fac, D3& brdf, D3& color,
           double& distFactor, D3& emis) {
    uint32_t ei = 0;
    const EmitterSample es = sample_emitter_idx(R.s, rng, ei);
    const V3 lightDir = normalize(es.pos - it.pos);
    const float dist = length(es.pos - it.pos);
    if (trace_occ(R.s, it.pos, lightDir, 1e-3f, dist)) return false;
    color = D3(R.s.evalEmit[ei]);
    const DMat m = resolve64(R.s, R.s.evalMats[R.s.triMat[it.prim]], it.prim, it.bu, it.bv);
    distFactor = 1.0 / ((double)dist * (double)dist);
    fac.mat = R.s.triMat[it.prim];
    fac.emit = ei;
    const float nDotV = dot(it.normal, -it.inDir), nDotL = dot(it.normal, lightDir);
    if (nDotV <= 0.0f || nDotL < 0.0f) {
        if (requireValue) return false;
        brdf = emis = D3(0.0);
        fac.v = fac.dA = fac.dR = fac.dC = D3(0.0);
        return true;
    }
    const ADouble b = brdf64(m, it.normal, -it.inDir, lightDir);
    brdf = b.val;
    emis = m.emissive;
    fac.v = b.val * color * distFactor;
    fac.dA = b.dA * color * distFactor;
    fac.dR = b.dR * color * distFactor;
    fac.dC = b.val * distFactor;
    return true;
}

AFactor bsdf_factor(AReplay& R, const Inter& it, const V3& outDir) {
    const ADouble b = brdf64(resolve64(R.s, R.s.evalMats[R.s.triMat[it.prim]], it.prim, it.bu, it.bv), it.normal,
                             -it.inDir, outDir);
    AFactor f;
    f.v = b.val;
    f.dA = b.dA;
    f.dR = b.dR;
    f.mat = R.s.triMat[it.prim];
    return f;
}

// TracePrefix (ReSTIR PT form) from vertex traceIdx0; see cx_adjoint.hip.
bool replay_tail64(AReplay& R, Inter it, bool haveIt, V3 pos, V3 dir, uint32_t traceIdx0, PCG32& rng, D3& prod,
                   D3& emis, bool& neeFirst) {
    const uint32_t maxLen = (uint32_t)R.P.lp.maxPathLen;
    prod = D3(1.0);
    emis = D3(0.0);
    neeFirst = false;
    for (uint32_t traceIdx = traceIdx0; traceIdx < maxLen; ++traceIdx) {
        if (!(haveIt && traceIdx == traceIdx0)) {
            trace_surface(R.s, pos, dir, 1e-3f, 1e16f, it);
            if (!it.valid) return false;
        }
        const bool forceNee = traceIdx == maxLen - 1;
        if (rng.next_float() < R.P.lp.neeProb || forceNee) {
            AFactor nf;
            D3 b, c, e;
            double df;
            if (!nee64(R, it, rng, true, nf, b, c, df, e)) return false;
            prod = prod * nf.v;
            R.F.push_back(nf);
            if (traceIdx == 0) {
                emis = e * df;
                neeFirst = true;
            }
            return true;
        }
        const BrdfSample bs = ggx_sample(R.s, it, rng);
        if (bs.samplingPdf <= 0.0f) return false;
        const AFactor f = bsdf_factor(R, it, bs.outDir);
        prod = prod * f.v;
        R.F.push_back(f);
        pos = it.pos;
        dir = bs.outDir;
    }
    return false;
}

Inter traced_inter(const Scene& s, const Packed& pk) {
    Inter r = rehydrate(s, pk);
    if (r.valid) r.normal = normalize(r.normal);
    return r;
}

bool replay_post_recon64(AReplay& R, const Prefix& src, D3& prc) {
    const uint32_t r = src.reconIdx(), L = src.length();
    PCG32 rng = src.rng;
    for (uint32_t i = 0; i < 4u * (r - 1u); i++) rng.next_u32();
    const Inter y = traced_inter(R.s, src.reconInt);
    if (!y.valid) return false;
    if (L == r) {
        rng.next_float();
        AFactor nf;
        D3 b, c, e;
        double df;
        if (!nee64(R, y, rng, true, nf, b, c, df, e)) return false;
        prc = nf.v;
        R.F.push_back(nf);
        return true;
    }
    rng.next_float();
    const BrdfSample bs = ggx_sample(R.s, y, rng);
    if (bs.samplingPdf <= 0.0f) return false;
    const AFactor yf = bsdf_factor(R, y, bs.outDir);
    R.F.push_back(yf);
    D3 tail, emis;
    bool first;
    Inter none;
    if (!replay_tail64(R, none, false, y.pos, bs.outDir, r, rng, tail, emis, first)) return false;
    prc = yf.v * tail;
    return true;
}

bool replay_shift64(AReplay& R, const Prefix& src, D3& f, D3& mask) {
    if (!src.valid() || src.reconIdx() < 2) return false;
    Inter curr = rehydrate(R.s, src.primaryInt);
    Inter srcRecon = rehydrate(R.s, src.reconInt);
    if (!curr.valid || !srcRecon.valid) return false;
    const int rvc = (int)src.reconIdx() - 2;
    PCG32 srng = src.rng;
    D3 thr(1.0);
    for (int i = 0; i < rvc; ++i) {
        srng.next_float();
        const BrdfSample bs = ggx_sample(R.s, curr, srng);
        if (bs.samplingPdf <= 0.0f) return false;
        const AFactor fb_ = bsdf_factor(R, curr, bs.outDir);
        thr = thr * fb_.v;
        R.F.push_back(fb_);
        trace_surface(R.s, curr.pos, bs.outDir, 1e-3f, 1e16f, curr);
        if (!curr.valid) return false;
    }
    const V3 reconVec = srcRecon.pos - curr.pos;
    const float reconLen = length(reconVec);
    const V3 reconDir = normalize(reconVec);
    if (std::isinf(reconLen) || any_nan_inf(reconDir)) return false;
    if (trace_occ(R.s, curr.pos, reconDir, 1e-3f, reconLen)) return false;
    if (dot(curr.normal, -curr.inDir) <= 0.0f || dot(curr.normal, reconDir) < 0.0f) return false;
    const AFactor b1 = bsdf_factor(R, curr, reconDir);
    thr = thr * b1.v;
    R.F.push_back(b1);
    srcRecon.inDir = reconDir;
    if (dot(srcRecon.normal, -srcRecon.inDir) <= 0.0f || dot(srcRecon.normal, src.reconOutDir) < 0.0f) return false;
    const AFactor b2 = bsdf_factor(R, srcRecon, src.reconOutDir);
    thr = thr * b2.v;
    R.F.push_back(b2);
    D3 prc;
    if (!replay_post_recon64(R, src, prc)) return false;
    const D3 x = thr * prc;
    f = D3(std::max(0.0, x.x), std::max(0.0, x.y), std::max(0.0, x.z));
    mask = D3(x.x > 0 ? 1.0 : 0.0, x.y > 0 ? 1.0 : 0.0, x.z > 0 ? 1.0 : 0.0);
    return true;
}

void contract64(const std::vector<AFactor>& F, const D3& u, std::vector<double>& gA, std::vector<double>& gR,
                std::vector<double>& gE) {
    const size_t n = F.size();
    std::vector<D3> suf(n + 1, D3(1.0));
    for (size_t k = n; k-- > 0;) suf[k] = F[k].v * suf[k + 1];
    D3 pre(1.0);
    for (size_t k = 0; k < n; ++k) {
        const AFactor& x = F[k];
        const D3 w = u * (pre * suf[k + 1]);
        const D3 ga = w * x.dA;
        gA[3 * x.mat + 0] += ga.x;
        gA[3 * x.mat + 1] += ga.y;
        gA[3 * x.mat + 2] += ga.z;
        gR[x.mat] += ddot(w, x.dR);
        if (x.emit != 0xFFFFFFFFu) {
            const D3 ge = w * x.dC;
            gE[3 * x.emit + 0] += ge.x;
            gE[3 * x.emit + 1] += ge.y;
            gE[3 * x.emit + 2] += ge.z;
        }
        pre = pre * x.v;
    }
}

// One pixel; returns the replayed estimate and accumulates into the thread's gradients.
// status: 0 replayed, 2 skipped, -1 nothing to replay.
int adjoint_pixel64(const Renderer& R, uint32_t p, const D3& g, D3& out, std::vector<double>& gA,
                    std::vector<double>& gR, std::vector<double>& gE) {
    const Scene& s = *R.scene;
    AReplay A{s, R.P, {}};
    out = D3(0.0);
    if (R.P.lp.rendererType == CX_RESTIR_PT) {
        const Reservoir<Prefix>& res = R.prefixRes[2 * p + R.P.front];
        if (!R.gbuf[p].primaryIntValid || !res.sample.valid() || res.wSum == 0.0f) return -1;
        PCG32 canon(sample_tea64(p, R.P.lp.random));
        canon.next2d();
        D3 f(0.0), mask(1.0);
        bool ok;
        if (res.sample.rng.state == canon.state) {
            PCG32 rng = res.sample.rng;
            D3 prod, emis;
            bool neeFirst = false;
            const Inter prim = traced_inter(s, res.sample.primaryInt);
            ok = prim.valid && replay_tail64(A, prim, true, V3(0.0f), V3(0.0f), 0u, rng, prod, emis, neeFirst);
            f = (neeFirst && res.sample.length() == 1) ? prod + emis : prod;
        } else {
            ok = replay_shift64(A, res.sample, f, mask);
        }
        if (!ok) return 2;
        const double W = res.wSum;
        out = f * W;
        contract64(A.F, g * mask * W, gA, gR, gE);
        return 0;
    }
    // path tracer
    const uint32_t x = p % R.W, y = p / R.W;
    PCG32 rng(sample_tea64(p, R.P.lp.random));
    V3 pos, dir;
    camera_ray(R, x, y, rng, pos, dir);
    const double neeProb = R.P.lp.neeProb;
    D3 thr(1.0), radiance(0.0);
    uint32_t depth = 0;
    bool contributes = false;
    for (uint32_t traceIdx = 0; traceIdx < (uint32_t)R.P.lp.maxPathLen; ++traceIdx) {
        Inter it;
        trace_surface(s, pos, dir, 1e-3f, 1e16f, it);
        if (!it.valid) break;
        ++depth;
        if (rng.next_float() < R.P.lp.neeProb) {
            thr = thr / neeProb;
            AFactor nf;
            D3 b, c, e;
            double df;
            if (!nee64(A, it, rng, false, nf, b, c, df, e)) break;
            radiance = thr * b * c * df;
            if (depth == 1) radiance = radiance + e;
            nf.v = nf.v / neeProb;
            nf.dA = nf.dA / neeProb;
            nf.dR = nf.dR / neeProb;
            nf.dC = nf.dC / neeProb;
            A.F.push_back(nf);
            contributes = true;
            break;
        }
        const BrdfSample bs = ggx_sample(s, it, rng);
        if (bs.samplingPdf <= 0.0f) break;
        AFactor f = bsdf_factor(A, it, bs.outDir);
        const double k = 1.0 / ((double)bs.samplingPdf * (1.0 - neeProb));
        f.v = f.v * k;
        f.dA = f.dA * k;
        f.dR = f.dR * k;
        thr = thr * f.v;
        A.F.push_back(f);
        pos = it.pos;
        dir = bs.outDir;
    }
    if (!contributes) return -1;
    out = radiance;
    contract64(A.F, g, gA, gR, gE);
    return 0;
}

}  // namespace

extern "C" {

int orc_abi_version(void) { return 1; }

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}


// FP64 truth image: spp samples per pixel of the PathTracer estimator through the pinhole
// camera (Camera.h:14-22 with a uniform jitter in the pixel).  out: [h][w][3] mean,
// var: [h][w][3] variance of the mean (may be null).
void orc_truth_render(const void* sp, const cx_camera* cam, uint32_t w, uint32_t h, uint32_t maxPathLen,
                      double neeProb, uint32_t spp, uint64_t seed, double* out, double* var) {
    const Scene& s = *(const Scene*)sp;
    const D3 cp(cam->pos[0], cam->pos[1], cam->pos[2]), U(cam->U[0], cam->U[1], cam->U[2]),
        V(cam->V[0], cam->V[1], cam->V[2]), W(cam->W[0], cam->W[1], cam->W[2]);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t p = 0; p < (int64_t)w * h; p++) {
        const uint32_t x = (uint32_t)(p % w), y = (uint32_t)(p / w);
        SplitMix rng{seed * 0x9E3779B97F4A7C15ULL + (uint64_t)p * 0xD1B54A32D192ED03ULL};
        rng.next();
        D3 sum(0.0), sq(0.0);
        for (uint32_t i = 0; i < spp; i++) {
            const double ux = 2.0 * ((x + rng.uniform()) / w) - 1.0, uy = 2.0 * ((y + rng.uniform()) / h) - 1.0;
            const D3 c = d_path(s, cp, dnorm(U * ux + V * uy + W), maxPathLen, neeProb, rng);
            sum = sum + c;
            sq = sq + c * c;
        }
        const D3 mean = sum / (double)spp;
        out[3 * p] = mean.x; out[3 * p + 1] = mean.y; out[3 * p + 2] = mean.z;
        if (var) {
            const D3 v = (sq / (double)spp - mean * mean) / (double)std::max(1u, spp - 1);
            var[3 * p] = v.x; var[3 * p + 1] = v.y; var[3 * p + 2] = v.z;
        }
    }
}

uint32_t orc_tea32(uint32_t a, uint32_t b, uint32_t rounds) { return sample_tea32(a, b, rounds); }
uint64_t orc_tea64(uint32_t a, uint32_t b, uint32_t rounds) { return sample_tea64(a, b, rounds); }
void orc_tea_keyed(uint32_t* v, const uint32_t* key, uint32_t rounds) { tea(v[0], v[1], key, rounds); }

// n draws of NextUint32 (kind 0), NextFloat bits (kind 1) or NextUint64 (kind 2, two u32 per value)
void orc_pcg32_stream(uint64_t initstate, uint64_t initseq, int kind, uint32_t n, uint32_t* out) {
    PCG32 r(initstate, initseq);
    for (uint32_t i = 0; i < n; i++) {
        if (kind == 0) out[i] = r.next_u32();
        else if (kind == 1) { const float f = r.next_float(); std::memcpy(&out[i], &f, 4); }
        else { const uint64_t v = r.next_u64(); out[2 * i] = (uint32_t)v; out[2 * i + 1] = (uint32_t)(v >> 32); }
    }
}

void orc_pcg32_state(uint64_t initstate, uint64_t initseq, uint64_t* stateInc) {
    PCG32 r(initstate, initseq);
    stateInc[0] = r.state;
    stateInc[1] = r.inc;
}

void orc_sincos(const float* x, uint32_t n, float* s, float* c) {
    for (uint32_t i = 0; i < n; i++) sincos_cw(x[i], &s[i], &c[i]);
}

void* orc_scene_create(const cx_scene_desc* d, int bruteForce) {
    Scene* s = new Scene();
    s->pos.resize(d->numVertices);
    s->nrm.resize(d->numVertices);
    for (uint32_t i = 0; i < d->numVertices; i++) {
        s->pos[i] = V3(d->positions[3 * i], d->positions[3 * i + 1], d->positions[3 * i + 2]);
        s->nrm[i] = V3(d->normals[3 * i], d->normals[3 * i + 1], d->normals[3 * i + 2]);
    }
    s->idx.assign(d->indices, d->indices + 3 * (size_t)d->numTriangles);
    s->triMat.assign(d->triMaterial, d->triMaterial + d->numTriangles);
    if (d->triMask) s->triVis.assign(d->triMask, d->triMask + d->numTriangles);
    else s->triVis.assign(d->numTriangles, (uint8_t)CX_VIS_ALL);
    s->mats.resize(d->numMaterials);
    for (uint32_t m = 0; m < d->numMaterials; m++) {
        const cx_material& cm = d->materials[m];
        s->mats[m].diff = V3(cm.diffColor[0], cm.diffColor[1], cm.diffColor[2]);
        s->mats[m].spec = V3(cm.specF0[0], cm.specF0[1], cm.specF0[2]);
        s->mats[m].emissive = V3(cm.emissiveColor[0], cm.emissiveColor[1], cm.emissiveColor[2]);
        s->mats[m].roughness = cm.roughness;
        if (d->materialTex) {
            const cx_material_tex& mt = d->materialTex[m];
            s->mats[m].diffTex = mt.diffTex;
            s->mats[m].specTex = mt.specTex;
            s->mats[m].roughTex = mt.roughTex;
            s->mats[m].specUsage = mt.specTexUsage;
        }
    }
    if (d->uvs) s->uv.assign(d->uvs, d->uvs + 2 * (size_t)d->numVertices);
    s->textures.resize(d->numTextures);
    for (uint32_t t = 0; t < d->numTextures; t++) {
        s->textures[t].w = d->textures[t].width;
        s->textures[t].h = d->textures[t].height;
        s->textures[t].rgba.assign(d->textures[t].rgba,
                                   d->textures[t].rgba + 4 * (size_t)d->textures[t].width * d->textures[t].height);
    }
    s->emitters.resize(d->numEmitters);
    for (uint32_t e = 0; e < d->numEmitters; e++) {
        EmitterData& ed = s->emitters[e];
        ed.color = V3(d->emitterColor[3 * e], d->emitterColor[3 * e + 1], d->emitterColor[3 * e + 2]);
        for (uint32_t f = d->emitterFaceOffset[e]; f < d->emitterFaceOffset[e + 1]; f++)
            for (int k = 0; k < 3; k++) {
                const float* v = &d->emitterFaces[(size_t)f * 9 + k * 3];
                ed.faces.push_back(V3(v[0], v[1], v[2]));
            }
    }
    // Scene::GetSceneSize (src/Scene.cpp:48-71): bbox of all vertices and the origin
    V3 mn(0.0f), mx(0.0f);
    for (const V3& p : s->pos) {
        mn = V3(gmin(mn.x, p.x), gmin(mn.y, p.y), gmin(mn.z, p.z));
        mx = V3(gmax(mx.x, p.x), gmax(mx.y, p.y), gmax(mx.z, p.z));
    }
    s->sceneSize = distance(mn, mx);
    s->bmin = mn; s->bmax = mx;
    s->evalMats = s->mats;
    for (const EmitterData& e : s->emitters) s->evalEmit.push_back(e.color);
    s->bruteForce = bruteForce != 0;
    build_bvh(*s);
    return s;
}

void orc_scene_destroy(void* s) { delete (Scene*)s; }
float orc_scene_size(void* s) { return ((Scene*)s)->sceneSize; }

// rays [n][8]; prim UINT32_MAX on miss
void orc_trace_closest(void* sp, const float* rays, uint32_t n, uint32_t* prim, float* t, float* u, float* v) {
    const Scene& s = *(Scene*)sp;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        const float* r = &rays[i * 8];
        const Hit h = trace_closest(s, V3(r[0], r[1], r[2]), V3(r[3], r[4], r[5]), r[6], r[7]);
        prim[i] = h.prim;
        t[i] = h.prim == INVALID_PRIM ? 0.0f : h.t;
        u[i] = h.prim == INVALID_PRIM ? 0.0f : h.u;
        v[i] = h.prim == INVALID_PRIM ? 0.0f : h.v;
    }
}

void orc_trace_occluded(void* sp, const float* rays, uint32_t n, uint32_t* occ) {
    const Scene& s = *(Scene*)sp;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        const float* r = &rays[i * 8];
        occ[i] = trace_occluded(s, V3(r[0], r[1], r[2]), V3(r[3], r[4], r[5]), r[6], r[7]) ? 1u : 0u;
    }
}

// BSDF probes.  inter: [n][8] = prim, bu, bv, inDir xyz, normal-mode (0 = traced/normalised,
// 1 = rehydrated/un-normalised), pad.  eval out: [n][8] = brdf xyz, pdf, emission xyz, roughness.
static Inter probe_inter(const Scene& s, const float* r) {
    Packed pk;
    uint32_t prim;
    std::memcpy(&prim, &r[0], 4);
    pk.prim = prim; pk.bu = r[1]; pk.bv = r[2]; pk.inDir = V3(r[3], r[4], r[5]); pk.valid = true; pk.hasMesh = true;
    Inter it = rehydrate(s, pk);
    if (r[6] == 0.0f) it.normal = normalize(it.normal);
    return it;
}

void orc_ggx_eval(void* sp, const float* inter, const float* outDir, uint32_t n, float* out) {
    const Scene& s = *(Scene*)sp;
    for (uint32_t i = 0; i < n; i++) {
        const Inter it = probe_inter(s, &inter[i * 8]);
        const BrdfEval e = ggx_eval(s, it, V3(outDir[3 * i], outDir[3 * i + 1], outDir[3 * i + 2]));
        v3out(e.brdfResult, &out[i * 8]);
        out[i * 8 + 3] = e.samplingPdf;
        v3out(e.emission, &out[i * 8 + 4]);
        out[i * 8 + 7] = e.roughness;
    }
}

// sample out: [n][8] = outDir xyz, brdfVal xyz, pdf, roughness; rng state per sample in/out
void orc_ggx_sample(void* sp, const float* inter, uint64_t* rngState, uint32_t n, float* out) {
    const Scene& s = *(Scene*)sp;
    for (uint32_t i = 0; i < n; i++) {
        const Inter it = probe_inter(s, &inter[i * 8]);
        PCG32 rng;
        rng.state = rngState[i];
        const BrdfSample b = ggx_sample(s, it, rng);
        rngState[i] = rng.state;
        v3out(b.outDir, &out[i * 8]);
        v3out(b.brdfVal, &out[i * 8 + 3]);
        out[i * 8 + 6] = b.samplingPdf;
        out[i * 8 + 7] = b.roughness;
    }
}

// emitter sample out: [n][4] = pos xyz, emitter index (as float); rng in/out
void orc_sample_emitter(void* sp, uint64_t* rngState, uint32_t n, float* out) {
    const Scene& s = *(Scene*)sp;
    for (uint32_t i = 0; i < n; i++) {
        PCG32 rng;
        rng.state = rngState[i];
        const EmitterSample e = sample_emitter(s, rng);
        rngState[i] = rng.state;
        v3out(e.pos, &out[i * 4]);
        out[i * 4 + 3] = e.color.x;
    }
}

float orc_jacobian(const float* oldXi, const float* newXi, const float* tp, const float* tn) {
    return reconnection_jacobian(V3(oldXi[0], oldXi[1], oldXi[2]), V3(newXi[0], newXi[1], newXi[2]),
                                 V3(tp[0], tp[1], tp[2]), V3(tn[0], tn[1], tn[2]));
}

void orc_pairwise_mis(const float* fBasis, const float* fNeigh, float J, float K, float canonM, float neighM,
                      float* canonW, float* neighW) {
    const V3 a(fBasis[0], fBasis[1], fBasis[2]), b(fNeigh[0], fNeigh[1], fNeigh[2]);
    *canonW = canonical_pairwise_mis(a, b, J, K, canonM, neighM);
    *neighW = neighbor_pairwise_mis(a, b, J, K, canonM, neighM);
}

// Reservoir known-answer sequences: streams n (f-luminance carrier, weight, confidence)
// candidates into one reservoir; returns the selected index, wSum and confidence before
// finalisation.
int orc_reservoir_stream(const float* weights, const float* confidences, uint32_t n, uint64_t rngState,
                         float* wSumOut, float* confOut) {
    struct Tagged { V3 f; int idx = -1; };
    Reservoir<Tagged> r;
    PCG32 rng;
    rng.state = rngState;
    for (uint32_t i = 0; i < n; i++) {
        Tagged t; t.f = V3(1.0f); t.idx = (int)i;
        r.update(t, weights[i], rng, confidences[i]);
    }
    *wSumOut = r.wSum;
    *confOut = r.confidence;
    return r.sample.idx;
}

// Identity shift probes: trace a prefix / suffix for pixel p of a fresh renderer-free setup
// and shift it onto itself.  Returns the shifted f and jacobian of the canonical prefix traced
// from (origin, dir) with the given rng state.  out: [8] = prefix f xyz, shifted f xyz, J, valid.
void orc_identity_shift_prefix(void* sp, const cx_launch_params* lp, const float* origin, const float* dir,
                               uint64_t rngState, float* out) {
    const Scene& s = *(Scene*)sp;
    Params P;
    P.lp = *lp;
    PCG32 rng;
    rng.state = rngState;
    Prefix pre;
    trace_prefix(s, pre, V3(origin[0], origin[1], origin[2]), V3(dir[0], dir[1], dir[2]), rng, P);
    float J = 0.0f;
    const V3 f = shift_prefix(s, pre, pre, J);
    v3out(pre.f, out);
    v3out(f, out + 3);
    out[6] = J;
    out[7] = (pre.valid() && pre.reconIdx() >= 2) ? 1.0f : 0.0f;
}

void orc_identity_shift_suffix(void* sp, const cx_launch_params* lp, const float* origin, const float* dir,
                               uint64_t rngState, float* out) {
    const Scene& s = *(Scene*)sp;
    Params P;
    P.lp = *lp;
    PCG32 rng;
    rng.state = rngState;
    Prefix pre;
    trace_prefix(s, pre, V3(origin[0], origin[1], origin[2]), V3(dir[0], dir[1], dir[2]), rng, P);
    Suffix suf;
    uint32_t reconVertex = 0;
    trace_suffix(s, suf, pre, rng, P, &reconVertex);
    float J = 0.0f;
    const V3 f = shift_suffix(s, suf, suf, J);
    v3out(suf.f, out);
    v3out(f, out + 3);
    out[6] = J;
    out[7] = suf.valid() ? 1.0f : 0.0f;
    out[8] = (float)suf.reconIdx();
    out[9] = (float)suf.length();
    out[10] = (float)reconVertex;                  // BSDF reconnection: the suffix vertex chosen, else 0
    out[11] = (float)pre.length();
}

// The final gather's combination (final_gather.cu:135-187) over injected neighbour terms, each
// 13 floats: fFromCanonOfNeigh[3], jNeighToCanon, neighF[3], neighWSum, neighConf,
// fFromNeighOfCanon[3], jCanonToNeigh.
void orc_gather_combine(const float* canonPreF, float canonPreP, const float* canonSufF, const float* canonContrib,
                        const float* terms, uint32_t n, float* out) {
    std::vector<GatherTerm> t(n);
    for (uint32_t i = 0; i < n; i++) {
        const float* x = terms + 13 * (size_t)i;
        t[i].fFromCanonOfNeigh = V3(x[0], x[1], x[2]);
        t[i].jNeighToCanon = x[3];
        t[i].neighF = V3(x[4], x[5], x[6]);
        t[i].neighWSum = x[7];
        t[i].neighConf = x[8];
        t[i].fFromNeighOfCanon = V3(x[9], x[10], x[11]);
        t[i].jCanonToNeigh = x[12];
    }
    const V3 r = gather_combine(V3(canonPreF[0], canonPreF[1], canonPreF[2]), canonPreP,
                                V3(canonSufF[0], canonSufF[1], canonSufF[2]),
                                V3(canonContrib[0], canonContrib[1], canonContrib[2]), t.data(), n);
    v3out(r, out);
}

void* orc_renderer_create(void* sp, uint32_t w, uint32_t h) {
    Renderer* R = new Renderer();
    R->scene = (Scene*)sp;
    R->W = w;
    R->H = h;
    std::memset(&R->P.lp, 0, sizeof(R->P.lp));
    renderer_reset(*R);
    R->frameRng = PCG32(0);
    return R;
}

void orc_renderer_destroy(void* r) { delete (Renderer*)r; }
void orc_renderer_set_params(void* r, const cx_launch_params* lp) { ((Renderer*)r)->P.lp = *lp; }
void orc_renderer_seed(void* r, uint64_t seed) { ((Renderer*)r)->frameRng = PCG32(seed); }
void orc_renderer_reset(void* r) { renderer_reset(*(Renderer*)r); }

void orc_renderer_frame(void* r, const cx_camera* cam, int camChanged, float* out) {
    launch_frame(*(Renderer*)r, *cam, camChanged != 0, out);
}

void orc_renderer_pass_ms(void* r, double* ms) {
    for (int i = 0; i < CX_PASS_COUNT; i++) ms[i] = ((Renderer*)r)->passMs[i];
}

// Closest-hit and shadow rays of each pass of the last frame: [2 * pass + {0, 1}].
void orc_renderer_rays(void* r, uint64_t* out) {
    for (int i = 0; i < 2 * CX_PASS_COUNT; i++) out[i] = ((Renderer*)r)->passRays[i];
}

void orc_renderer_buffers(void* r, uint32_t* frontBack, uint64_t* frameIdxRandom) {
    const Renderer& R = *(Renderer*)r;
    frontBack[0] = R.P.front;
    frontBack[1] = R.P.back;
    frameIdxRandom[0] = R.frameIdx;
    frameIdxRandom[1] = R.P.lp.random;
}

void orc_dump_prefix(void* r, int which, orc_prefix_dump* out) {
    const Renderer& R = *(Renderer*)r;
    for (uint32_t p = 0; p < R.pix(); p++) {
        const Reservoir<Prefix>& res = R.prefixRes[2 * p + which];
        const Prefix& s = res.sample;
        orc_prefix_dump& d = out[p];
        d.rngState = s.rng.state; d.flags = s.flags;
        v3out(s.f, d.f); v3out(s.postReconF, d.postReconF);
        d.p = s.p; d.pathLen = s.pathLen; v3out(s.reconOutDir, d.reconOutDir);
        d.wSum = res.wSum; d.confidence = res.confidence;
        dump_packed(s.primaryInt, d.primaryInt); dump_packed(s.reconInt, d.reconInt); dump_packed(s.lastInt, d.lastInt);
    }
}

void orc_dump_suffix(void* r, int which, orc_suffix_dump* out) {
    const Renderer& R = *(Renderer*)r;
    for (uint32_t p = 0; p < R.pix(); p++) {
        const Reservoir<Suffix>& res = R.suffixRes[2 * p + which];
        const Suffix& s = res.sample;
        orc_suffix_dump& d = out[p];
        d.rngState = s.rng.state; d.flags = s.flags;
        v3out(s.f, d.f); v3out(s.postReconF, d.postReconF);
        d.p = s.p; v3out(s.reconOutDir, d.reconOutDir);
        d.wSum = res.wSum; d.confidence = res.confidence;
        dump_packed(s.reconInt, d.reconInt); dump_packed(s.lastPrefixInt, d.lastPrefixInt);
    }
}

// gbuffer: [P][4] u32 = rng lo, rng hi, primaryIntValid, canon flags (bit0 canon prefix valid,
// bit1 canon primaryInt valid, bit2 canon suffix valid)
void orc_dump_gbuffer(void* r, uint32_t* out) {
    const Renderer& R = *(Renderer*)r;
    for (uint32_t p = 0; p < R.pix(); p++) {
        out[p * 4 + 0] = (uint32_t)R.gbuf[p].rng.state;
        out[p * 4 + 1] = (uint32_t)(R.gbuf[p].rng.state >> 32);
        out[p * 4 + 2] = R.gbuf[p].primaryIntValid ? 1u : 0u;
        out[p * 4 + 3] = (R.canonPrefix[p].valid() ? 1u : 0u) | (R.canonPrefix[p].primaryInt.valid ? 2u : 0u) |
                         (R.canonSuffix[p].valid() ? 4u : 0u);
    }
}

// entries: pos [P][3], radius [P], valid [P]; neighbours [P][k] and counts [P]
void orc_dump_entries(void* r, float* pos, float* radius, uint8_t* valid) {
    const Renderer& R = *(Renderer*)r;
    for (uint32_t p = 0; p < R.pix(); p++) {
        valid[p] = R.entryValid[p];
        radius[p] = (R.entryBox[p * 6 + 3] - R.entryBox[p * 6 + 0]) / 2.0f;
        const Inter last = rehydrate(*R.scene, R.prefixRes[2 * p + R.P.front].sample.lastInt);
        v3out(R.entryValid[p] ? last.pos : V3(0.0f), &pos[p * 3]);
    }
}

uint32_t orc_dump_neighbors(void* r, uint32_t* idx, float* dist, uint32_t* count) {
    const Renderer& R = *(Renderer*)r;
    const uint32_t k = (uint32_t)std::max(0, R.P.lp.gatherM - 1);
    for (uint32_t p = 0; p < R.pix(); p++) {
        count[p] = R.neighCount.empty() ? 0 : R.neighCount[p];
        for (uint32_t i = 0; i < k; i++) {
            const Neighbor& nb = R.neighbors[(size_t)p * k + i];
            idx[(size_t)p * k + i] = i < count[p] ? nb.pixelIdx : 0xFFFFFFFFu;
            dist[(size_t)p * k + i] = i < count[p] ? nb.distance : 0.0f;
        }
    }
    return k;
}

// Standalone neighbour query over an entry set (radius = requested r, put through the box
// round trip of prefix_store_entries.cu:43-48 / prefix_search_hit.cu:33-34).
void orc_entries_query(const float* pos, const float* r, const uint8_t* valid, uint32_t n, const float* q,
                       uint32_t nq, uint32_t k, int brute, uint32_t* idx, float* dist, uint32_t* count) {
    EntrySet E;
    E.pos.resize(n);
    E.radius.resize(n);
    E.valid.assign(valid, valid + n);
    for (uint32_t i = 0; i < n; i++) {
        E.pos[i] = V3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        const float lo = pos[3 * i] - r[i], hi = pos[3 * i] + r[i];
        E.radius[i] = (hi - lo) / 2.0f;
    }
    entries_build(E);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t j = 0; j < (int64_t)nq; j++) {
        std::vector<Neighbor> nb(std::max(1u, k));
        const V3 qq(q[3 * j], q[3 * j + 1], q[3 * j + 2]);
        const uint32_t c = brute ? entries_query_brute(E, qq, k, nb.data()) : entries_query(E, qq, k, nb.data());
        count[j] = c;
        for (uint32_t i = 0; i < k; i++) {
            idx[(size_t)j * k + i] = i < c ? nb[i].pixelIdx : 0xFFFFFFFFu;
            dist[(size_t)j * k + i] = i < c ? nb[i].distance : 0.0f;
        }
    }
}


// FP64 path-replay adjoint of the last frame (see above).  dLdI: [P][3] (may be null: replay
// only); gA [M][3], gR [M], gE [E][3] are overwritten (null to skip); replay [P][3] (may be
// null); status[3]: pixels replayed, -, pixels skipped.
void orc_renderer_adjoint(void* rp, const float* dLdI, double* gA, double* gR, double* gE, double* replay,
                          uint32_t* status) {
    const Renderer& R = *(const Renderer*)rp;
    const Scene& s = *R.scene;
    const size_t M = s.mats.size(), E = s.emitters.size();
    const int64_t n = R.pix();
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    std::vector<std::vector<double>> tA(nt, std::vector<double>(3 * M, 0.0)), tR(nt, std::vector<double>(M, 0.0)),
        tE(nt, std::vector<double>(3 * E, 0.0));
    std::vector<uint32_t> st(3 * (size_t)nt, 0u);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t pi = 0; pi < n; pi++) {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        const D3 g = dLdI ? D3(dLdI[3 * pi], dLdI[3 * pi + 1], dLdI[3 * pi + 2]) : D3(0.0);
        D3 out;
        const int rc = adjoint_pixel64(R, (uint32_t)pi, g, out, tA[t], tR[t], tE[t]);
        if (rc == 0) st[3 * t + 0]++;
        if (rc == 2) st[3 * t + 2]++;
        if (replay) {
            replay[3 * pi + 0] = out.x;
            replay[3 * pi + 1] = out.y;
            replay[3 * pi + 2] = out.z;
        }
    }
    for (int t = 0; t < nt; t++) {
        for (size_t i = 0; i < 3 * M; i++)
            if (gA) gA[i] = (t ? gA[i] : 0.0) + tA[t][i];
        for (size_t i = 0; i < M; i++)
            if (gR) gR[i] = (t ? gR[i] : 0.0) + tR[t][i];
        for (size_t i = 0; i < 3 * E; i++)
            if (gE) gE[i] = (t ? gE[i] : 0.0) + tE[t][i];
        if (status)
            for (int k = 0; k < 3; k++) status[k] = (t ? status[k] : 0u) + st[3 * t + k];
    }
}

// Evaluation (2) and sampling (1) tables of the oracle scene.
void orc_scene_set_materials(void* sp, const cx_material* mats, uint32_t n, int tables) {
    Scene& s = *(Scene*)sp;
    for (uint32_t m = 0; m < n && m < s.mats.size(); m++) {
        Material x = s.mats[m];   // texture bindings stay
        x.diff = V3(mats[m].diffColor[0], mats[m].diffColor[1], mats[m].diffColor[2]);
        x.spec = V3(mats[m].specF0[0], mats[m].specF0[1], mats[m].specF0[2]);
        x.emissive = V3(mats[m].emissiveColor[0], mats[m].emissiveColor[1], mats[m].emissiveColor[2]);
        x.roughness = mats[m].roughness;
        if (tables & 1) s.mats[m] = x;
        if (tables & 2) s.evalMats[m] = x;
    }
}

// Samples one texture at n (u, v) pairs with the canonical filter; out [n][3].
void orc_texture_sample(const cx_texture* tx, const float* uv, uint32_t n, float* out) {
    Texture t;
    t.w = tx->width;
    t.h = tx->height;
    t.rgba.assign(tx->rgba, tx->rgba + 4 * (size_t)t.w * t.h);
    for (uint32_t i = 0; i < n; i++) v3out(tex_fetch(t, uv[2 * i], uv[2 * i + 1]), &out[3 * i]);
}

void orc_scene_set_emitter_colors(void* sp, const float* colors, uint32_t n, int tables) {
    Scene& s = *(Scene*)sp;
    for (uint32_t e = 0; e < n && e < s.emitters.size(); e++) {
        const V3 c(colors[3 * e], colors[3 * e + 1], colors[3 * e + 2]);
        if (tables & 1) s.emitters[e].color = c;
        if (tables & 2) s.evalEmit[e] = c;
    }
}


// LinearToSrgb / ApplySrgbGamma (src/kernel/tonemap.cu:7-24): the sRGB curve selected as
// glm::lerp does, glm's clamp to [0, 1], x 255, truncation to a byte (NaN -> 0).  The power is
// evaluated in double and rounded once to float.
void orc_tonemap(const float* hdr, uint8_t* ldr, uint32_t n) {
    for (uint32_t i = 0; i < n; i++) {
        const float c = hdr[i];
        const bool cond = c < 0.0031308f;
        const float ifTrue = 12.92f * c;
        const float p = (float)std::pow((double)c, (double)(1.0f / 2.4f));
        const float ifFalse = 1.055f * p - 0.055f;
        const float t = cond ? 1.0f : 0.0f;
        float v = ifFalse * (1.0f - t) + ifTrue * t;
        v = (v < 0.0f) ? 0.0f : v;
        v = (1.0f < v) ? 1.0f : v;
        v = v * 255.0f;
        ldr[i] = std::isnan(v) ? (uint8_t)0 : (uint8_t)v;
    }
}

}  // extern "C"
