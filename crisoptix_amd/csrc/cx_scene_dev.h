// Scene access on the device: BVH traversal (closest hit and occlusion), hit records,
// rehydration of packed interactions, the GGX BSDF and emitter sampling.
//
// Traversal replaces optixTrace over the triangle GAS/IAS (include/graph/trace.h:25-88).
// The result of a query is a pure function of the ray and the triangle set:
//   closest hit  = the minimum (t, primitive) over all triangles whose watertight test
//                  (Woop, Benthin, Wald, JCGT 2013) reports t in [tmin, tmax];
//   occlusion    = whether any such triangle exists.
// Node boxes are padded outward at build time and the box test carries a relative slack,
// so traversal only ever visits more nodes than needed; which BVH is used, and in which
// order children are visited, cannot change a result.
#pragma once

#include "cx_dev.h"

namespace cx {

static constexpr uint32_t INVALID_PRIM = 0xFFFFFFFFu;
static constexpr float PI_F = 3.14159265358979323846264f;   // shader/brdf.cu:12

// Binary BVH node, 64 bytes:
//   a = (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y)   b = (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y)
//   c = (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z)   d = (ref0, ref1, -, -)
// ref >= 0: inner node index; ref < 0: leaf, ~ref = (first triangle slot << 4) | count.
// Triangle slots hold 3 float4: (v0.xyz, primitive index bits), (v1.xyz, skip bits), (v2.xyz, 0);
// the skip bits are the complement of the triangle's CX_VIS_* mask: bit 0 hides it from
// closest-hit rays, bit 1 from occlusion rays.
struct DMaterial {
    float4 diffRough;   // diffuse rgb, roughness
    float4 spec;        // specF0 rgb, -
    float4 emissive;    // emissive rgb, -
    int4 tex;           // diffuse, specular, roughness texture (-1: none), specular usage
};

static constexpr int32_t SPEC_USAGE_OCC_ROUGH_METAL = 1;   // SpecTexUsage::OccRoughMetal

// Wide BVH node (width N = 4 or 8), N/4 groups of 7 float4, padded to 8 float4 per group:
//   lo.x[4], hi.x[4], lo.y[4], hi.y[4], lo.z[4], hi.z[4], ref[4]
// with refs as in the binary node; an empty slot has ref WIDE_EMPTY (never a leaf code, since
// a leaf's first slot plus its count stays below 2^27) and a point box at WIDE_EMPTY_COORD,
// which no ray with tmax <= 1e30 can reach.
static constexpr int32_t WIDE_EMPTY = (int32_t)0x80000000;
static constexpr float WIDE_EMPTY_COORD = 3.0e38f;

struct DScene {
    const float4* __restrict__ nodes;
    const float4* __restrict__ wnodes;    // wide nodes (width 4 or 8), or null
    const float4* __restrict__ tris;
    const float* __restrict__ pos;        // [V][3]
    const float* __restrict__ nrm;        // [V][3]
    const uint32_t* __restrict__ idx;     // [T][3]
    const uint32_t* __restrict__ triMat;  // [T]
    const DMaterial* __restrict__ mats;
    const float* __restrict__ emitFaces;  // [F][9]
    const uint32_t* __restrict__ emitOff; // [E+1]
    const float* __restrict__ emitColor;  // [E][3]
    uint32_t numTris;
    uint32_t numEmitters;
    int32_t root;         // root of `nodes` (binary)
    int32_t wroot;        // root of `wnodes`
    uint32_t width;       // 2: traverse `nodes`; 4 or 8: traverse `wnodes`
    uint32_t numTextures;
    uint32_t pad[2];
    const float* __restrict__ uv;         // [V][2], or null (all zero)
    const uint32_t* __restrict__ texels;  // RGBA8 texels of every texture, one word each
    const uint4* __restrict__ texInfo;    // [textures]: first texel, width, height, -
};

static constexpr int STACK_DEPTH = 64;

struct RayPre {
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

CX_HD RayPre ray_pre(V3 d) {
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    RayPre r;
    r.kz = (ax >= ay) ? ((ax >= az) ? 0 : 2) : ((ay >= az) ? 1 : 2);
    r.kx = r.kz == 2 ? 0 : r.kz + 1;
    r.ky = r.kx == 2 ? 0 : r.kx + 1;
    if (comp(d, r.kz) < 0.0f) {
        const int t = r.kx;
        r.kx = r.ky;
        r.ky = t;
    }
    const float dz = comp(d, r.kz);
    r.Sx = comp(d, r.kx) / dz;
    r.Sy = comp(d, r.ky) / dz;
    r.Sz = 1.0f / dz;
    return r;
}

// Watertight ray/triangle test.  u weights v1 and v weights v2 (optixGetTriangleBarycentrics).
CX_HD bool tri_hit(V3 o, const RayPre& r, V3 v0, V3 v1, V3 v2, float tmin, float tmax,
                                        float& t, float& u, float& v) {
    const V3 A = v0 - o, B = v1 - o, C = v2 - o;
    const float Akx = comp(A, r.kx), Aky = comp(A, r.ky), Akz = comp(A, r.kz);
    const float Bkx = comp(B, r.kx), Bky = comp(B, r.ky), Bkz = comp(B, r.kz);
    const float Ckx = comp(C, r.kx), Cky = comp(C, r.ky), Ckz = comp(C, r.kz);
    const float Ax = Akx - r.Sx * Akz, Ay = Aky - r.Sy * Akz;
    const float Bx = Bkx - r.Sx * Bkz, By = Bky - r.Sy * Bkz;
    const float Cx = Ckx - r.Sx * Ckz, Cy = Cky - r.Sy * Ckz;
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
    const float Az = r.Sz * Akz, Bz = r.Sz * Bkz, Cz = r.Sz * Ckz;
    const float T = U * Az + V * Bz + W * Cz;
    const float tt = T / det;
    if (!(tt >= tmin && tt <= tmax)) return false;
    t = tt;
    u = V / det;
    v = W / det;
    return true;
}

struct Hit {
    uint32_t prim;
    float t, u, v;
};

// Slab test of both children of a node; the far plane carries a relative slack so that a
// triangle whose computed t equals the current best is never pruned.
CX_HD void box2(const float4 a, const float4 b, const float4 c, V3 o, V3 inv, float tmin,
                                     float tcap, bool& hit0, bool& hit1, float& tn0, float& tn1) {
    const float l0x = (a.x - o.x) * inv.x, h0x = (a.y - o.x) * inv.x;
    const float l0y = (a.z - o.y) * inv.y, h0y = (a.w - o.y) * inv.y;
    const float l0z = (c.x - o.z) * inv.z, h0z = (c.y - o.z) * inv.z;
    const float l1x = (b.x - o.x) * inv.x, h1x = (b.y - o.x) * inv.x;
    const float l1y = (b.z - o.y) * inv.y, h1y = (b.w - o.y) * inv.y;
    const float l1z = (c.z - o.z) * inv.z, h1z = (c.w - o.z) * inv.z;
    const float n0 = fmaxf(fmaxf(fminf(l0x, h0x), fminf(l0y, h0y)), fmaxf(fminf(l0z, h0z), tmin));
    const float f0 = fminf(fminf(fmaxf(l0x, h0x), fmaxf(l0y, h0y)), fminf(fmaxf(l0z, h0z), tcap));
    const float n1 = fmaxf(fmaxf(fminf(l1x, h1x), fminf(l1y, h1y)), fmaxf(fminf(l1z, h1z), tmin));
    const float f1 = fminf(fminf(fmaxf(l1x, h1x), fmaxf(l1y, h1y)), fminf(fmaxf(l1z, h1z), tcap));
    hit0 = n0 <= f0 * 1.00002f + 1e-30f;
    hit1 = n1 <= f1 * 1.00002f + 1e-30f;
    tn0 = n0;
    tn1 = n1;
}

__device__ __forceinline__ V3 load_v3(const float* p, uint32_t i) { return V3{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }

// Traversal stack: the first SHORT entries in a per-thread LDS column (lds[i * stride]), deeper
// entries in a per-lane private array.
template <int SHORT>
struct TravStack {
    int* lds;
    int stride;
    int ovf[STACK_DEPTH - SHORT > 0 ? STACK_DEPTH - SHORT : 1];
    CX_HD void put(int i, int v) {
        if (SHORT >= STACK_DEPTH || i < SHORT) lds[i * stride] = v;
        else ovf[i - SHORT] = v;
    }
    CX_HD int get(int i) const {
        if (SHORT >= STACK_DEPTH || i < SHORT) return lds[i * stride];
        return ovf[i - SHORT];
    }
};

// A traversal stack with its top entry in a register: a pop takes the register,
// and the stack read that refills it is needed only at the next push or pop, so it is not on
// the path from a pop to the fetch of the popped node.
template <typename Stack>
struct TopStack {
    Stack& st;
    int& top;
    // push of entry i (= the current depth)
    CX_HD void put(int i, int v) {
        if (i > 0) st.put(i - 1, top);
        top = v;
    }
};

// Closest-hit / any-hit leaf loop shared by the binary and wide traversals.  Returns true when
// an any-hit query is answered.  Triangles whose skip bits hide them from the ray kind (closest
// hit: bit 0, any hit: bit 1) are passed over, as OptiX passes over an instance whose
// visibilityMask shares no bit with the ray's.  The interval tested is [tmin, best.t]: best.t
// starts at the ray's tmax and only falls, so a triangle behind the current closest hit is
// rejected before its barycentric divisions, and one at exactly best.t still reaches the
// (t, primitive) tie-break.
CX_HD bool leaf_hits_rt(const DScene& S, int32_t ref, V3 o, const RayPre& rp, float tmin, Hit& best, bool any) {
    const uint32_t code = ~(uint32_t)ref;
    const uint32_t first = code >> 4, cnt = code & 15u;
    // the next triangle's vertices are fetched before the current one is tested, so a leaf's
    // fetches overlap instead of queueing behind each test (a leaf holds at least one triangle)
    const float4* tv = S.tris + 3 * (size_t)first;
    float4 n0 = tv[0], n1 = tv[1], n2 = tv[2];
    for (uint32_t i = 0; i < cnt; i++) {
        const float4 p0 = n0, p1 = n1, p2 = n2;
        if (i + 1 < cnt) {
            const float4* nv = tv + 3 * (size_t)(i + 1);
            n0 = nv[0];
            n1 = nv[1];
            n2 = nv[2];
        }
        if (__builtin_bit_cast(uint32_t, p1.w) & (any ? 2u : 1u)) continue;   // invisible to this ray kind
        float t, u, v;
        if (!tri_hit(o, rp, V3{p0.x, p0.y, p0.z}, V3{p1.x, p1.y, p1.z}, V3{p2.x, p2.y, p2.z}, tmin, best.t, t, u, v))
            continue;
        const uint32_t prim = __builtin_bit_cast(uint32_t, p0.w);
        if (any) {
            best.prim = prim;
            best.t = t;
            return true;
        }
        if (t < best.t || (t == best.t && prim < best.prim)) {
            best.prim = prim;
            best.t = t;
            best.u = u;
            best.v = v;
        }
    }
    return false;
}

template <bool ANY>
CX_HD bool leaf_hits(const DScene& S, int32_t ref, V3 o, const RayPre& rp, float tmin, Hit& best) {
    return leaf_hits_rt(S, ref, o, rp, tmin, best, ANY);
}

// One triangle of a leaf, its slot already loaded in q[0..2]: leaf_hits' test of that triangle.
// Returns true when an any-hit query is answered.
CX_HD bool tri_hits_q(const float4* q, V3 o, const RayPre& rp, float tmin, Hit& best, bool any) {
    if (__builtin_bit_cast(uint32_t, q[1].w) & (any ? 2u : 1u)) return false;   // invisible to this ray kind
    float t, u, v;
    if (!tri_hit(o, rp, V3{q[0].x, q[0].y, q[0].z}, V3{q[1].x, q[1].y, q[1].z}, V3{q[2].x, q[2].y, q[2].z}, tmin,
                 best.t, t, u, v))
        return false;
    const uint32_t prim = __builtin_bit_cast(uint32_t, q[0].w);
    if (any) {
        best.prim = prim;
        best.t = t;
        return true;
    }
    if (t < best.t || (t == best.t && prim < best.prim)) {
        best.prim = prim;
        best.t = t;
        best.u = u;
        best.v = v;
    }
    return false;
}

// Slab test of one child of a wide node (same arithmetic and slack as box2).
CX_HD float box1(float lx, float hx, float ly, float hy, float lz, float hz, V3 o, V3 inv,
                                      float tmin, float tcap, bool& hit) {
    const float ax = (lx - o.x) * inv.x, bx = (hx - o.x) * inv.x;
    const float ay = (ly - o.y) * inv.y, by = (hy - o.y) * inv.y;
    const float az = (lz - o.z) * inv.z, bz = (hz - o.z) * inv.z;
    const float n = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), tmin));
    const float f = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), tcap));
    hit = n <= f * 1.00002f + 1e-30f;
    return n;
}

CX_HD void cswap(float& ta, int& ra, float& tb, int& rb) {
    const bool sw = tb < ta;
    const float t = sw ? tb : ta;
    const int r = sw ? rb : ra;
    tb = sw ? ta : tb;
    rb = sw ? ra : rb;
    ta = t;
    ra = r;
}

// One wide node: test the N child boxes, sort the hit ones by entry distance (a fixed
// compare-exchange network on registers, misses at +inf), push all but the nearest, farthest
// first, and return the nearest in ref; false (ref unchanged) when no child box is hit.  q holds
// the node's first 7 float4 per group of 4 children (bounds lo.x, hi.x, lo.y, hi.y, lo.z, hi.z
// of the 4 children, then their references), already loaded.
template <int N, typename Stack>
CX_HD bool wide_visit_q(const float4* q, int& ref, V3 o, V3 inv, float tmin, float tcap, Stack& st, int& sp) {
    float tn[N];
    int cr[N];
#pragma unroll
    for (int g = 0; g < N / 4; g++) {
        const float4 lx = q[7 * g + 0], hx = q[7 * g + 1], ly = q[7 * g + 2], hy = q[7 * g + 3];
        const float4 lz = q[7 * g + 4], hz = q[7 * g + 5];
        const int4 rr = __builtin_bit_cast(int4, q[7 * g + 6]);
        const float lxa[4] = {lx.x, lx.y, lx.z, lx.w}, hxa[4] = {hx.x, hx.y, hx.z, hx.w};
        const float lya[4] = {ly.x, ly.y, ly.z, ly.w}, hya[4] = {hy.x, hy.y, hy.z, hy.w};
        const float lza[4] = {lz.x, lz.y, lz.z, lz.w}, hza[4] = {hz.x, hz.y, hz.z, hz.w};
        const int ra[4] = {rr.x, rr.y, rr.z, rr.w};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            bool h;
            const float t = box1(lxa[j], hxa[j], lya[j], hya[j], lza[j], hza[j], o, inv, tmin, tcap, h);
            const bool use = h && ra[j] != WIDE_EMPTY;
            tn[4 * g + j] = use ? t : INFINITY;
            cr[4 * g + j] = ra[j];
        }
    }
    if constexpr (N == 4) {
        cswap(tn[0], cr[0], tn[1], cr[1]);
        cswap(tn[2], cr[2], tn[3], cr[3]);
        cswap(tn[0], cr[0], tn[2], cr[2]);
        cswap(tn[1], cr[1], tn[3], cr[3]);
        cswap(tn[1], cr[1], tn[2], cr[2]);
    } else {
        // Batcher's odd-even merge sort of 8 keys (19 compare-exchanges).
#define CX_CS(a, b) cswap(tn[a], cr[a], tn[b], cr[b])
        CX_CS(0, 1); CX_CS(2, 3); CX_CS(4, 5); CX_CS(6, 7);
        CX_CS(0, 2); CX_CS(1, 3); CX_CS(4, 6); CX_CS(5, 7);
        CX_CS(1, 2); CX_CS(5, 6);
        CX_CS(0, 4); CX_CS(1, 5); CX_CS(2, 6); CX_CS(3, 7);
        CX_CS(2, 4); CX_CS(3, 5);
        CX_CS(1, 2); CX_CS(3, 4); CX_CS(5, 6);
#undef CX_CS
    }
    if (tn[0] == INFINITY) return false;
#pragma unroll
    for (int i = N - 1; i >= 1; i--) {
        if (tn[i] != INFINITY) {
            st.put(sp, cr[i]);
            ++sp;
        }
    }
    ref = cr[0];
    return true;
}

// wide_visit_q on node `ref` of the wide array (2 N float4 per node, the 8th of each group of
// 4 children unused).
template <int N, typename Stack>
CX_HD bool wide_visit(const DScene& S, int& ref, V3 o, V3 inv, float tmin, float tcap, Stack& st, int& sp) {
    const float4* n = S.wnodes + (size_t)ref * (2 * N);
    float4 q[7 * (N / 4)];
#pragma unroll
    for (int g = 0; g < N / 4; g++) {
#pragma unroll
        for (int k = 0; k < 7; k++) q[7 * g + k] = n[8 * g + k];
    }
    return wide_visit_q<N>(q, ref, o, inv, tmin, tcap, st, sp);
}

// One binary node: the nearer hit child in ref (the farther one pushed); false when neither
// child box is hit.
template <typename Stack>
CX_HD bool bin_visit(const DScene& S, int& ref, V3 o, V3 inv, float tmin, float tcap, Stack& st, int& sp) {
    const float4* n = S.nodes + 4 * (size_t)ref;
    const float4 a = n[0], b = n[1], c = n[2];
    const int4 dd = __builtin_bit_cast(int4, n[3]);
    bool h0, h1;
    float t0, t1;
    box2(a, b, c, o, inv, tmin, tcap, h0, h1, t0, t1);
    if (h0 && h1) {
        const bool first0 = t0 <= t1;
        st.put(sp, first0 ? dd.y : dd.x);
        ++sp;
        ref = first0 ? dd.x : dd.y;
        return true;
    }
    if (h0) { ref = dd.x; return true; }
    if (h1) { ref = dd.y; return true; }
    return false;
}

template <bool ANY, int N, typename Stack>
CX_HD bool traverse_wide(const DScene& S, V3 o, V3 d, float tmin, float tmax, Hit& best,
                                              Stack& st) {
    const RayPre rp = ray_pre(d);
    const V3 inv = V3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    int sp = 0;
    int ref = S.wroot;
    while (true) {
        if (ref >= 0) {
            if (wide_visit<N>(S, ref, o, inv, tmin, best.t, st, sp)) continue;
        } else {
            if (leaf_hits<ANY>(S, ref, o, rp, tmin, best)) return true;
        }
        if (sp == 0) break;
        --sp;
        ref = st.get(sp);
    }
    return best.prim != INVALID_PRIM;
}

template <bool ANY, typename Stack>
CX_HD bool traverse_bin(const DScene& S, V3 o, V3 d, float tmin, float tmax, Hit& best,
                                             Stack& st) {
    best.prim = INVALID_PRIM;
    best.t = tmax;
    best.u = 0.0f;
    best.v = 0.0f;
    if (S.numTris == 0) return false;
    const RayPre rp = ray_pre(d);
    const V3 inv = V3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    int sp = 0;
    int ref = S.root;
    while (true) {
        if (ref >= 0) {
            if (bin_visit(S, ref, o, inv, tmin, best.t, st, sp)) continue;
        } else {
            if (leaf_hits<ANY>(S, ref, o, rp, tmin, best)) return true;
        }
        if (sp == 0) break;
        --sp;
        ref = st.get(sp);
    }
    return best.prim != INVALID_PRIM;
}

// The traversal loop above as a resumable state, advanced one node or one leaf per step, for
// kernels that interleave several rays per lane (a lane whose ray ends takes the next one
// without waiting for the rest of its wavefront).  The answer is the loop's: the same nodes in
// the same order.
struct TravRay {
    V3 o, inv;
    RayPre rp;
    float tmin;
    Hit best;   // best.t: the ray's tmax until a hit lowers it
    int ref;
    int sp;
    int top;   // stack entry sp - 1 (valid when sp > 0); the stack holds entries 0 .. sp - 2
};


template <int WIDTH>
CX_HD void trav_init(const DScene& S, TravRay& r, V3 o, V3 d, float tmin, float tmax) {
    r.o = o;
    r.inv = V3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    r.rp = ray_pre(d);
    r.tmin = tmin;
    r.best = Hit{INVALID_PRIM, tmax, 0.0f, 0.0f};
    r.ref = WIDTH == 2 ? S.root : S.wroot;
    r.sp = 0;
    r.top = 0;
}

// One step; true when the ray is answered: closest hit in r.best (prim INVALID_PRIM: a miss),
// occlusion as r.best.prim != INVALID_PRIM.  A scene without triangles answers at once.
//
// At width 4 a step makes one fetch for both kinds of lane: a lane at an inner node loads the
// node's 7 float4, a lane at a leaf the 7 float4 from the slot of its next triangle (that
// triangle's 3, then 4 unused), in the same load instructions, before either kind is processed.
// A leaf step tests one triangle and leaves the lane on the rest of the leaf, so the leaf branch
// of a wavefront split between nodes and leaves costs one triangle test per step.  A wavefront whose
// lanes are split between nodes and leaves so waits once per step instead of once for the node
// fetches and again for the triangle fetches.
template <int WIDTH, typename Stack>
CX_HD bool trav_step_rt(const DScene& S, TravRay& r, Stack& stack, bool any) {
    if (S.numTris == 0) return true;
    TopStack<Stack> st{stack, r.top};
    if constexpr (WIDTH == 4) {
        const bool inner = r.ref >= 0;
        const uint32_t code = ~(uint32_t)r.ref;
        // a leaf's 7 loads run past its next triangle into the following slots, or into the
        // TRI_PAD float4 after the last slot; one base address, seven immediate offsets
        const float4* base = inner ? S.wnodes + (size_t)r.ref * 8 : S.tris + 3 * (size_t)(code >> 4);
        float4 q[7];
#pragma unroll
        for (uint32_t k = 0; k < 7; k++) q[k] = base[k];
        if (inner) {
            if (wide_visit_q<4>(q, r.ref, r.o, r.inv, r.tmin, r.best.t, st, r.sp)) return false;
        } else {
            if (tri_hits_q(q, r.o, r.rp, r.tmin, r.best, any)) return true;
            if ((code & 15u) > 1u) {
                r.ref = (int32_t)~(code + 15u);   // the rest of the leaf, from its next slot
                return false;
            }
        }
    } else if (r.ref >= 0) {
        if constexpr (WIDTH == 2) {
            if (bin_visit(S, r.ref, r.o, r.inv, r.tmin, r.best.t, st, r.sp)) return false;
        } else {
            if (wide_visit<WIDTH>(S, r.ref, r.o, r.inv, r.tmin, r.best.t, st, r.sp)) return false;
        }
    } else {
        if (leaf_hits_rt(S, r.ref, r.o, r.rp, r.tmin, r.best, any)) return true;
    }
    if (r.sp == 0) return true;
    --r.sp;
    r.ref = r.top;
    if (r.sp > 0) r.top = stack.get(r.sp - 1);
    return false;
}

template <bool ANY, int WIDTH, typename Stack>
CX_HD bool trav_step(const DScene& S, TravRay& r, Stack& st) {
    return trav_step_rt<WIDTH>(S, r, st, ANY);
}

// Traversal at a BVH width fixed at compile time (WIDTH = S.width; the caller dispatches).
template <bool ANY, int WIDTH, typename Stack>
CX_HD bool traverse_w(const DScene& S, V3 o, V3 d, float tmin, float tmax, Hit& best, Stack& st) {
    best.prim = INVALID_PRIM;
    best.t = tmax;
    best.u = 0.0f;
    best.v = 0.0f;
    if (S.numTris == 0) return false;
    if constexpr (WIDTH == 2) return traverse_bin<ANY>(S, o, d, tmin, tmax, best, st);
    else return traverse_wide<ANY, WIDTH>(S, o, d, tmin, tmax, best, st);
}

// Dispatch on the scene's BVH width (uniform across a launch).
template <bool ANY, typename Stack>
CX_HD bool traverse_s(const DScene& S, V3 o, V3 d, float tmin, float tmax, Hit& best, Stack& st) {
    if (S.width == 4) return traverse_w<ANY, 4>(S, o, d, tmin, tmax, best, st);
    if (S.width == 8) return traverse_w<ANY, 8>(S, o, d, tmin, tmax, best, st);
    return traverse_w<ANY, 2>(S, o, d, tmin, tmax, best, st);
}

// Traversal with the whole stack in a per-thread LDS column (stack[i * stride]).
template <bool ANY>
__device__ __forceinline__ bool traverse(const DScene& S, V3 o, V3 d, float tmin, float tmax, Hit& best,
                                         int* stack, int stride) {
    TravStack<STACK_DEPTH> st;
    st.lds = stack;
    st.stride = stride;
    return traverse_s<ANY>(S, o, d, tmin, tmax, best, st);
}

// ---- interactions (include/graph/Interaction.h) ----

// PackedInteraction, one flattened mesh (instance 0): prim == INVALID_PRIM means !valid.
struct Packed {
    uint32_t prim;
    float bu, bv;
    V3 dir;
};

CX_HD Packed packed_invalid() { return Packed{INVALID_PRIM, 0.0f, 0.0f, V3{0.0f, 0.0f, 0.0f}}; }

struct Inter {
    uint32_t prim;
    float bu, bv;
    V3 pos, dir, nrm;
    bool valid;
    CX_HD Packed pack() const {
        Packed p;
        p.prim = valid ? prim : INVALID_PRIM;
        p.bu = bu;
        p.bv = bv;
        p.dir = dir;
        return p;
    }
};

// InterpolateBary (Interaction.h:18-25)
CX_HD V3 interp(float u, float v, V3 a, V3 b, V3 c) { return ((1.0f - u - v) * a + u * b) + v * c; }

// Rehydration (Interaction.h:115-172): normals are interpolated and not normalised.
__device__ __forceinline__ Inter rehydrate(const DScene& S, const Packed& p) {
    Inter r;
    r.prim = p.prim;
    r.bu = p.bu;
    r.bv = p.bv;
    r.dir = p.dir;
    r.valid = p.prim != INVALID_PRIM && p.prim < S.numTris;
    r.pos = v3(0.0f);
    r.nrm = v3(0.0f);
    if (!r.valid) return r;
    const uint32_t i0 = S.idx[3 * p.prim], i1 = S.idx[3 * p.prim + 1], i2 = S.idx[3 * p.prim + 2];
    r.pos = interp(p.bu, p.bv, load_v3(S.pos, i0), load_v3(S.pos, i1), load_v3(S.pos, i2));
    r.nrm = interp(p.bu, p.bv, load_v3(S.nrm, i0), load_v3(S.nrm, i1), load_v3(S.nrm, i2));
    return r;
}

// Per-thread traversal context and ray counters.
struct TraceCtx {
    int* stack;
    int stride;
    uint32_t closest;
    uint32_t shadow;
};

// TraceWithDataPointer<Interaction> + __closesthit__mesh / __miss__main
// (trace.h:25-45, mesh_hit.cu:17-69, miss.cu:11-19): a miss keeps every other field.
__device__ __forceinline__ void trace_surface(const DScene& S, TraceCtx& tc, V3 o, V3 d, float tmin, float tmax,
                                              Inter& io) {
    Hit h;
    tc.closest++;
    traverse<false>(S, o, d, tmin, tmax, h, tc.stack, tc.stride);
    io.dir = d;
    if (h.prim == INVALID_PRIM) {
        io.valid = false;
        return;
    }
    io.valid = true;
    io.prim = h.prim;
    io.bu = h.u;
    io.bv = h.v;
    const uint32_t i0 = S.idx[3 * h.prim], i1 = S.idx[3 * h.prim + 1], i2 = S.idx[3 * h.prim + 2];
    io.pos = interp(h.u, h.v, load_v3(S.pos, i0), load_v3(S.pos, i1), load_v3(S.pos, i2));
    io.nrm = normalize(interp(h.u, h.v, load_v3(S.nrm, i0), load_v3(S.nrm, i1), load_v3(S.nrm, i2)));
}

// TraceOcclusion (trace.h:63-88)
__device__ __forceinline__ bool trace_occluded(const DScene& S, TraceCtx& tc, V3 o, V3 d, float tmin, float tmax) {
    Hit h;
    tc.shadow++;
    return traverse<true>(S, o, d, tmin, tmax, h, tc.stack, tc.stride);
}

// ---- BSDF (shader/brdf.cu) ----

struct BrdfEval {
    V3 brdf;
    float pdf;
    V3 emission;
    float roughness;
};

struct BrdfSample {
    V3 outDir, brdf;
    float pdf;
    float roughness;
};

CX_HD float d_ggx(float NdotH, float roughness) {   // brdf.cu:99-104
    const float a2 = roughness * roughness;
    const float d = (NdotH * a2 - NdotH) * NdotH + 1.0f;
    return a2 / (PI_F * d * d);
}

CX_HD float v_smith(float NdotL, float NdotV, float roughness) {   // brdf.cu:106-112
    const float a2 = roughness * roughness;
    const float lv = NdotL * sqrtf(NdotV * NdotV * (1 - a2) + a2);
    const float ll = NdotV * sqrtf(NdotL * NdotL * (1 - a2) + a2);
    return 0.5f / (lv + ll);
}

struct Mat {
    V3 diff, spec, emis;
    float rough;
    float metal;
    bool diffParam;    // diff is the material's constant (not a texture value)
    bool roughParam;   // rough is the material's constant
};

// tex2D<float4> of an RGBA8 texture created as in src/Texture.cpp:44-76 (wrap addressing,
// linear filtering, normalized coordinates, normalized-float reads), channels r, g, b.
// Canonical filter: x = frac(u) * width - 0.5, texels floor(x) and floor(x) + 1 wrapped,
// weight alpha = frac(x) rounded to a multiple of 1/256 (likewise y); the exact integer sum of
// weight * texel is rounded once, sum / (65536 * 255).  frac(u) = u - floor(u) lies in [0, 1],
// so floor(x) lies in [-1, width - 1] for width <= 65536; a non-finite coordinate reads as 0.
__device__ __forceinline__ V3 tex_fetch(const DScene& S, int32_t t, float u, float v) {
    const uint4 ti = S.texInfo[t];
    if (!(fabsf(u) <= 3.0e38f)) u = 0.0f;
    if (!(fabsf(v) <= 3.0e38f)) v = 0.0f;
    const float x = (u - floorf(u)) * (float)ti.y - 0.5f;
    const float y = (v - floorf(v)) * (float)ti.z - 0.5f;
    const float xf = floorf(x), yf = floorf(y);
    const uint32_t a = (uint32_t)rintf((x - xf) * 256.0f), b = (uint32_t)rintf((y - yf) * 256.0f);
    const uint32_t i0 = xf < 0.0f ? ti.y - 1u : (uint32_t)xf, j0 = yf < 0.0f ? ti.z - 1u : (uint32_t)yf;
    const uint32_t i1 = i0 + 1u == ti.y ? 0u : i0 + 1u, j1 = j0 + 1u == ti.z ? 0u : j0 + 1u;
    const uint32_t* row0 = S.texels + ti.x + (size_t)j0 * ti.y;
    const uint32_t* row1 = S.texels + ti.x + (size_t)j1 * ti.y;
    const uint32_t t00 = row0[i0], t10 = row0[i1], t01 = row1[i0], t11 = row1[i1];
    const uint32_t w00 = (256u - a) * (256u - b), w10 = a * (256u - b), w01 = (256u - a) * b, w11 = a * b;
    float c[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const uint32_t sh = 8u * (uint32_t)k;
        const uint32_t sum = w00 * ((t00 >> sh) & 255u) + w10 * ((t10 >> sh) & 255u) + w01 * ((t01 >> sh) & 255u) +
                             w11 * ((t11 >> sh) & 255u);
        c[k] = (float)sum / 16711680.0f;
    }
    return V3{c[0], c[1], c[2]};
}

// The material of a hit: GetFromTexIfPossible for diffuse colour, specular value and roughness
// at the interaction's uv (InterpolateBary of the vertex uvs, Interaction.h:171), the
// OccRoughMetal channel mapping, then specF0 = mix(diffColor, specF0, metal) (brdf.cu:121-138).
__device__ __forceinline__ Mat resolve_mat(const DScene& S, const DMaterial& m, uint32_t prim, float bu, float bv) {
    Mat r;
    r.diff = V3{m.diffRough.x, m.diffRough.y, m.diffRough.z};
    r.rough = m.diffRough.w;
    V3 spec = V3{m.spec.x, m.spec.y, m.spec.z};
    r.emis = V3{m.emissive.x, m.emissive.y, m.emissive.z};
    r.diffParam = true;
    r.roughParam = true;
    float metal = 0.0f;
    const int4 t = m.tex;
    if ((t.x & t.y & t.z) != -1 || t.w == SPEC_USAGE_OCC_ROUGH_METAL) {
        float tu = 0.0f, tv = 0.0f;
        if ((t.x & t.y & t.z) != -1 && S.uv != nullptr) {
            const uint32_t i0 = S.idx[3 * prim], i1 = S.idx[3 * prim + 1], i2 = S.idx[3 * prim + 2];
            const float w = 1.0f - bu - bv;
            tu = (w * S.uv[2 * i0] + bu * S.uv[2 * i1]) + bv * S.uv[2 * i2];
            tv = (w * S.uv[2 * i0 + 1] + bu * S.uv[2 * i1 + 1]) + bv * S.uv[2 * i2 + 1];
        }
        if (t.x >= 0) {
            r.diff = tex_fetch(S, t.x, tu, tv);
            r.diffParam = false;
        }
        if (t.y >= 0) spec = tex_fetch(S, t.y, tu, tv);
        if (t.z >= 0) {
            r.rough = tex_fetch(S, t.z, tu, tv).x;
            r.roughParam = false;
        }
        if (t.w == SPEC_USAGE_OCC_ROUGH_METAL) {
            r.rough = spec.y;
            metal = spec.z;
            spec = v3(1.0f);
            r.roughParam = false;
        }
    }
    r.spec = r.diff * (1.0f - metal) + spec * metal;
    r.metal = metal;
    return r;
}

__device__ __forceinline__ Mat load_mat(const DScene& S, const Inter& it) {
    return resolve_mat(S, S.mats[S.triMat[it.prim]], it.prim, it.bu, it.bv);
}

__device__ __forceinline__ V3 spec_term(const Mat& m, float nDotH, float nDotL, float nDotV, float lDotH) {
    V3 specBrdf = v3(0.0f);
    if ((double)dot(m.spec, m.spec) > 1e-6) {
        const float d = d_ggx(nDotH, m.rough);
        const float v = v_smith(nDotL, nDotV, m.rough);
        const float p = pow5(gmax(0.0f, 1.0f - lDotH));
        const V3 f = m.spec + (v3(1.0f) - m.spec) * p;
        specBrdf = d * v * f;
    }
    return specBrdf;
}

// __direct_callable__ggx_eval (brdf.cu:115-190)
__device__ __forceinline__ BrdfEval ggx_eval(const DScene& S, const Inter& it, V3 outDir) {
    const Mat m = load_mat(S, it);
    BrdfEval r;
    const V3 lightDir = outDir;
    const V3 viewDir = -it.dir;
    const V3 normal = it.nrm;
    const float nDotV = dot(normal, viewDir);
    const float nDotL = dot(normal, lightDir);
    if (nDotV <= 0.0f || nDotL < 0.0f) {
        r.brdf = v3(0.0f);
        r.emission = v3(0.0f);
        r.pdf = 0.0f;
        r.roughness = 0.0f;
        return r;
    }
    const V3 diffBrdf = m.diff / PI_F;
    const float diffProb = dot(m.diff, v3(1.0f)) / (dot(m.diff, v3(1.0f)) + dot(m.spec, v3(1.0f)));
    const V3 halfway = normalize(lightDir + viewDir);
    const float nDotH = dot(halfway, normal);
    const float lDotH = dot(halfway, lightDir);
    const V3 specBrdf = spec_term(m, nDotH, nDotL, nDotV, lDotH);
    r.brdf = (diffBrdf + specBrdf) * nDotL;
    r.pdf = diffProb * nDotL / PI_F + (1.0f - diffProb) * d_ggx(nDotH, m.rough) * nDotH / (4.0f * lDotH);
    r.emission = m.emis;
    r.roughness = m.rough;
    return r;
}

// ComputeLocalFrame (brdf.cu:192-212) applied to a local vector: mat3 columns (X, Y, Z).
CX_HD V3 frame_mul(V3 n, V3 l) {
    const float x = n.x, y = n.y, z = n.z;
    const float sz = (z >= 0) ? 1.0f : -1.0f;
    const float a = 1 / (sz + z);
    const float ya = y * a;
    const float b = x * ya;
    const float c = x * sz;
    const V3 X = V3{c * x * a - 1, sz * b, c};
    const V3 Y = V3{b, y * ya - sz, y};
    return V3{X.x * l.x + Y.x * l.y + n.x * l.z, X.y * l.x + Y.y * l.y + n.y * l.z, X.z * l.x + Y.z * l.y + n.z * l.z};
}

// __direct_callable__ggx_sample (brdf.cu:214-338)
__device__ __forceinline__ BrdfSample ggx_sample(const DScene& S, const Inter& it, Rng& rng) {
    const Mat m = load_mat(S, it);
    const V3 normal = it.nrm;
    const V3 viewDir = -it.dir;
    BrdfSample r;
    r.roughness = m.rough;
    const float nDotV = dot(normal, viewDir);
    if (nDotV <= 0) {
        r.outDir = v3(0.0f);
        r.brdf = v3(0.0f);
        r.pdf = 0.0f;
        return r;
    }
    const float diffProb = dot(m.diff, v3(1.0f)) / (dot(m.diff, v3(1.0f)) + dot(m.spec, v3(1.0f)));
    if (rng.f() < diffProb) {
        const float phi = rng.f() * 2.0f * PI_F;
        const float rr = sqrtf(rng.f());
        float sp, cp;
        sincos_cw(phi, &sp, &cp);
        const float x = rr * cp;
        const float y = rr * sp;
        const float z = sqrtf(gclamp(1.0f - x * x - y * y, 0.0f, 1.0f));
        r.outDir = frame_mul(normal, V3{x, y, z});
    } else {
        const float u = rng.f();
        const float mz = sqrtf(gclamp((1 - u) / (1 + (m.rough * m.rough - 1) * u), 0.0f, 1.0f));
        const float mr = sqrtf(gclamp(1 - mz * mz, 0.0f, 1.0f));
        const float phi = rng.f() * 2.0f * PI_F;
        float sp, cp;
        sincos_cw(phi, &sp, &cp);
        r.outDir = reflect(it.dir, frame_mul(normal, V3{mr * cp, mr * sp, mz}));
        if (dot(r.outDir, normal) < 0.0f) {
            r.outDir = v3(0.0f);
            r.brdf = v3(0.0f);
            r.pdf = 0.0f;
            return r;
        }
    }
    const V3 diffBrdf = m.diff / PI_F;
    const V3 lightDir = r.outDir;
    const float nDotL = dot(normal, lightDir);
    const V3 halfway = normalize(lightDir + viewDir);
    const float nDotH = dot(halfway, normal);
    const float lDotH = dot(halfway, lightDir);
    const V3 specBrdf = spec_term(m, nDotH, nDotL, nDotV, lDotH);
    r.pdf = diffProb * nDotL / PI_F + (1.0f - diffProb) * d_ggx(nDotH, m.rough) * nDotH / (4.0f * lDotH);
    r.brdf = (diffBrdf + specBrdf) * nDotL;
    return r;
}

// ---- emitters (include/graph/sample_emitter.h:3-12, include/model/Emitter.h:21-54) ----

struct EmitterSample {
    V3 pos, color;
};

__device__ __forceinline__ EmitterSample sample_emitter(const DScene& S, Rng& rng) {
    const uint64_t ei = rng.u64() % (uint64_t)S.numEmitters;
    const uint32_t f0 = S.emitOff[ei], f1 = S.emitOff[ei + 1];
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
    EmitterSample e;
    e.pos = r0 * v0 + r1 * v1 + r2 * v2;
    e.color = V3{S.emitColor[3 * ei], S.emitColor[3 * ei + 1], S.emitColor[3 * ei + 2]};
    return e;
}

}  // namespace cx
