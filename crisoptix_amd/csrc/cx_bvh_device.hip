// BVH build on the device: PLOC clustering of Morton-ordered triangles into a binary tree,
// SAH leaf formation, and the collapse to 4- or 8-wide nodes, written in the node formats of
// cx_scene_dev.h (the same the host build writes).  It stands in for optixAccelBuild of the
// triangle GAS (src/Mesh.cpp:54-135, src/Model.cpp:376-457) as the host build does; the
// traversal's answers do not depend on the tree (minimum (t, primitive) over the watertight
// test with conservative boxes), so either build gives the same hits.
//
// Clustering follows Meister and Bittner, "Parallel Locally-Ordered Clustering for Bounding
// Volume Hierarchy Construction" (IEEE TVCG 24(12), 2018): triangles are sorted by the 30-bit
// Morton code of their box centroid over the centroid bounds (stable, so ties keep the triangle
// order); then, until one cluster is left, every cluster finds the
// cluster within PLOC_RADIUS positions of it whose merged box has the smallest surface area,
// mutual nearest neighbours merge into a new inner node (at the lower position), and the
// cluster array is compacted in order.  Ties in area are broken by the pair's positions: nearer
// pairs first, then pairs whose lower position is even, then the lower position.  That is a total
// order on pairs, so the globally best pair is always mutual and every iteration merges at least
// once; and a run of clusters with equal boxes (duplicated geometry) merges pairwise, (0,1),
// (2,3), ..., so it becomes a balanced subtree instead of a chain.
//
// Children are created in an earlier iteration than their parent, so:
//   * the bottom-up SAH decision runs inside node creation: cost(leaf triangle) = area;
//     cost(node) = min(area * triangles  [if triangles <= maxLeaf], 2 * area + cost(l) + cost(r)),
//     the host build's traversal-to-intersection cost ratio of 2;
//   * the top-down pass (triangle slots, kept inner nodes, depths) runs one launch per
//     iteration in reverse creation order, with no inter-thread synchronisation inside a kernel.
// Every step is deterministic (scans, no ordering by atomics).  Child boxes are padded outward by
// |coord| * 2^-15 when the node arrays are written, as in the host build.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "cx_internal.h"

namespace cx {

namespace {

constexpr int BB = 256;                   // threads per block
constexpr int SCAN_ITEMS = 4;             // elements per thread of the scan kernels
constexpr uint32_t SCAN_TILE = BB * SCAN_ITEMS;
// Cluster positions searched on each side.  Measured on the 1.06 M-triangle object grid (4-wide
// traversal of a 1080p Conditional ReSTIR frame): radius 16 / 32 / 64 -> 6.79 / 6.41 / 6.32 ms;
// a 60-bit Morton code instead of 30 bits moved it by 0.1-0.15 ms, up at one radius and down at
// the other, so the single 30-bit sort is kept.
constexpr int PLOC_RADIUS = 64;
constexpr uint32_t NODE_KEPT = 1, NODE_LEAF = 2, NODE_BELOW = 3;   // top-down states

__device__ __forceinline__ uint32_t tid() { return blockIdx.x * blockDim.x + threadIdx.x; }

__device__ __forceinline__ float box_area(float4 lo, float4 hi) {
    const float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    if (!(dx >= 0.0f)) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

__device__ __forceinline__ float4 min4(float4 a, float4 b) {
    return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z), 0.0f);
}
__device__ __forceinline__ float4 max4(float4 a, float4 b) {
    return make_float4(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z), 0.0f);
}

__device__ __forceinline__ float pad_of(float lo, float hi) {
    return fmaxf(fabsf(lo), fabsf(hi)) * 3.0517578125e-5f + 1e-20f;
}

// ---- scans: exclusive prefix sums of uint2 pairs, in place, over three launches ----

__device__ __forceinline__ uint2 add2(uint2 a, uint2 b) { return make_uint2(a.x + b.x, a.y + b.y); }

// Exclusive scan of one pair per thread over the block; total receives the block's sum.
__device__ uint2 block_scan2(uint2 v, uint2& total) {
    __shared__ uint2 s_w[BB / 64];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint2 x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t yx = __shfl_up(x.x, off, 64), yy = __shfl_up(x.y, off, 64);
        if (lane >= (uint32_t)off) {
            x.x += yx;
            x.y += yy;
        }
    }
    if (lane == 63u) s_w[w] = x;
    __syncthreads();
    uint2 before = make_uint2(0u, 0u);
    total = make_uint2(0u, 0u);
#pragma unroll
    for (int k = 0; k < BB / 64; k++) {
        const uint2 c = s_w[k];
        if ((uint32_t)k < w) before = add2(before, c);
        total = add2(total, c);
    }
    __syncthreads();
    return make_uint2(before.x + x.x - v.x, before.y + x.y - v.y);
}

// The element count is n, or *nDev when nDev is not null (the grid then covers an upper bound).
__global__ __launch_bounds__(BB) void k_scan_tiles(const uint2* a, uint32_t n, const uint32_t* nDev,
                                                   uint2* tileSums) {
    if (nDev) n = *nDev;
    const uint32_t t0 = blockIdx.x * SCAN_TILE;
    uint2 s = make_uint2(0u, 0u);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        const uint32_t k = t0 + (uint32_t)i * BB + threadIdx.x;
        if (k < n) s = add2(s, a[k]);
    }
    uint2 total;
    (void)block_scan2(s, total);
    if (threadIdx.x == 0) tileSums[blockIdx.x] = total;
}

__global__ __launch_bounds__(BB) void k_scan_top(uint2* tileSums, uint32_t tiles, const uint32_t* nDev,
                                                 uint2* total) {
    if (nDev) tiles = (*nDev + SCAN_TILE - 1) / SCAN_TILE;
    uint2 carry = make_uint2(0u, 0u);
    for (uint32_t b0 = 0; b0 < tiles; b0 += BB) {
        const uint32_t b = b0 + threadIdx.x;
        const uint2 v = b < tiles ? tileSums[b] : make_uint2(0u, 0u);
        uint2 t;
        const uint2 ex = block_scan2(v, t);
        if (b < tiles) tileSums[b] = add2(carry, ex);
        carry = add2(carry, t);
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(BB) void k_scan_apply(uint2* a, uint32_t n, const uint32_t* nDev,
                                                   const uint2* tileSums) {
    if (nDev) n = *nDev;
    if (blockIdx.x * SCAN_TILE >= n) return;   // a tile past the count (uniform per block)
    const uint32_t t0 = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint2 v[SCAN_ITEMS];
    uint2 s = make_uint2(0u, 0u);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        v[i] = t0 + i < n ? a[t0 + i] : make_uint2(0u, 0u);
        s = add2(s, v[i]);
    }
    uint2 total;
    uint2 run = add2(tileSums[blockIdx.x], block_scan2(s, total));
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        if (t0 + i < n) a[t0 + i] = run;
        run = add2(run, v[i]);
    }
}

// ---- triangle boxes, centroid bounds, Morton keys ----

__global__ __launch_bounds__(BB) void k_tri_boxes(const float* pos, const uint32_t* idx, uint32_t n, float4* triLo,
                                                  float4* triHi, float* partial) {
    __shared__ float s_r[6][BB / 64];
    const uint32_t t = tid();
    float c[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    if (t < n) {
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int k = 0; k < 3; k++) {
            const float* v = pos + 3 * (size_t)idx[3 * (size_t)t + k];
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], v[a]);
                hi[a] = std::max(hi[a], v[a]);
            }
        }
        triLo[t] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        triHi[t] = make_float4(hi[0], hi[1], hi[2], 0.0f);
        for (int a = 0; a < 3; a++) {
            const float m = 0.5f * (lo[a] + hi[a]);
            c[a] = m;
            c[3 + a] = m;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        for (int a = 0; a < 3; a++) {
            c[a] = fminf(c[a], __shfl_xor(c[a], off, 64));
            c[3 + a] = fmaxf(c[3 + a], __shfl_xor(c[3 + a], off, 64));
        }
    if ((threadIdx.x & 63u) == 0)
        for (int a = 0; a < 6; a++) s_r[a][threadIdx.x >> 6] = c[a];
    __syncthreads();
    if (threadIdx.x < 6) {
        float r = s_r[threadIdx.x][0];
        for (int w = 1; w < BB / 64; w++)
            r = threadIdx.x < 3 ? fminf(r, s_r[threadIdx.x][w]) : fmaxf(r, s_r[threadIdx.x][w]);
        partial[6 * (size_t)blockIdx.x + threadIdx.x] = r;
    }
}

__global__ __launch_bounds__(BB) void k_bounds(const float* partial, uint32_t blocks, float* bounds) {
    __shared__ float s_r[6][BB];
    float c[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint32_t b = threadIdx.x; b < blocks; b += BB)
        for (int a = 0; a < 3; a++) {
            c[a] = fminf(c[a], partial[6 * (size_t)b + a]);
            c[3 + a] = fmaxf(c[3 + a], partial[6 * (size_t)b + 3 + a]);
        }
    for (int a = 0; a < 6; a++) s_r[a][threadIdx.x] = c[a];
    __syncthreads();
    if (threadIdx.x < 6) {
        float r = s_r[threadIdx.x][0];
        for (int k = 1; k < BB; k++) r = threadIdx.x < 3 ? fminf(r, s_r[threadIdx.x][k]) : fmaxf(r, s_r[threadIdx.x][k]);
        bounds[threadIdx.x] = r;
    }
}

// Bit b of x to bit 3b (x < 2^10).
__device__ __forceinline__ uint32_t spread10(uint32_t x) {
    x &= 0x3FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// 30-bit Morton code of triangle t's box centroid (below 2^30, so never the sort's invalid key).
__global__ __launch_bounds__(BB) void k_morton(const float4* triLo, const float4* triHi, uint32_t n, const float* bounds,
                                               uint32_t* keys) {
    const uint32_t t = tid();
    if (t >= n) return;
    const float4 lo = triLo[t], hi = triHi[t];
    const float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
    uint32_t q[3];
    for (int a = 0; a < 3; a++) {
        const float ext = bounds[3 + a] - bounds[a];
        const float f = ext > 0.0f ? (c[a] - bounds[a]) * (1024.0f / ext) : 0.0f;
        q[a] = (uint32_t)std::min(1023.0f, std::max(0.0f, f));
    }
    keys[t] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
}

// ---- PLOC ----

// Leaf s (sorted position) holds triangle vals[s]; one cluster per leaf.
__global__ __launch_bounds__(BB) void k_ploc_init(const uint32_t* vals, const float4* triLo, const float4* triHi,
                                                  uint32_t n, float4* nlo, float4* nhi, uint32_t* cnt, float* cost,
                                                  uint32_t* leafOK, uint32_t* clusters) {
    const uint32_t s = tid();
    if (s >= n) return;
    const uint32_t t = vals[s];
    const float4 lo = triLo[t], hi = triHi[t];
    nlo[s] = lo;
    nhi[s] = hi;
    cnt[s] = 1u;
    cost[s] = box_area(lo, hi);
    leafOK[s] = 1u;
    clusters[s] = s;
}

// nn[i] = the position j != i within PLOC_RADIUS whose merged box with i has the smallest
// area, ties to the smaller (|i - j|, min(i, j) odd, min(i, j)); i itself when there is none.
// Device state of the clustering: the current cluster count, the inner nodes made so far, the
// iterations run, and a flag for an iteration that merged nothing.
struct PlocState {
    uint32_t m, nodeCount, iters, stalled;
};

__global__ __launch_bounds__(BB) void k_ploc_nn(const uint32_t* clusters, const PlocState* st, const float4* nlo,
                                                const float4* nhi, uint32_t* nn) {
    const uint32_t m = st->m;
    if (m <= 1 || blockIdx.x * BB >= m) return;   // finished, or a block past the clusters
    constexpr int W = BB + 2 * PLOC_RADIUS;
    __shared__ float4 s_lo[W], s_hi[W];
    const int base = (int)(blockIdx.x * BB) - PLOC_RADIUS;
    for (int k = threadIdx.x; k < W; k += BB) {
        const int j = base + k;
        if (j >= 0 && j < (int)m) {
            const uint32_t c = clusters[j];
            s_lo[k] = nlo[c];
            s_hi[k] = nhi[c];
        }
    }
    __syncthreads();
    const int i = (int)(blockIdx.x * BB + threadIdx.x);
    if (i >= (int)m) return;
    const float4 lo = s_lo[threadIdx.x + PLOC_RADIUS], hi = s_hi[threadIdx.x + PLOC_RADIUS];
    float best = INFINITY;
    unsigned long long bestKey = ~0ull;
    int bestJ = i;
    for (int d = -PLOC_RADIUS; d <= PLOC_RADIUS; d++) {
        const int j = i + d;
        if (d == 0 || j < 0 || j >= (int)m) continue;
        const int k = threadIdx.x + PLOC_RADIUS + d;
        const float a = box_area(min4(lo, s_lo[k]), max4(hi, s_hi[k]));
        const uint32_t lo = (uint32_t)min(i, j);
        const unsigned long long key = ((unsigned long long)(uint32_t)abs(d) << 33) |
                                       ((unsigned long long)(lo & 1u) << 32) | (unsigned long long)lo;
        if (a < best || (a == best && key < bestKey)) {
            best = a;
            bestKey = key;
            bestJ = j;
        }
    }
    nn[i] = (uint32_t)bestJ;
}

// flags[i] = (cluster i survives, cluster i creates a node): mutual nearest neighbours merge
// at the lower position.
__global__ __launch_bounds__(BB) void k_ploc_flags(const uint32_t* nn, const PlocState* st, uint2* flags) {
    const uint32_t m = st->m, i = tid();
    if (m <= 1 || i >= m) return;
    const uint32_t j = nn[i];
    const bool mutual = j != i && nn[j] == i;
    flags[i] = make_uint2(mutual && i > j ? 0u : 1u, mutual && i < j ? 1u : 0u);
}

// After the scan: flags[i] = (output position, rank among this iteration's new nodes).
__global__ __launch_bounds__(BB) void k_ploc_merge(const uint32_t* cin, uint32_t* cout, const uint32_t* nn,
                                                   const uint2* flags, const PlocState* st, uint32_t n,
                                                   uint32_t maxLeaf, float4* nlo, float4* nhi, uint2* kids,
                                                   uint32_t* cnt, float* cost, uint32_t* leafOK) {
    const uint32_t m = st->m, nodeBase = st->nodeCount, i = tid();
    if (m <= 1 || i >= m) return;
    const uint32_t j = nn[i];
    const bool mutual = j != i && nn[j] == i;
    if (mutual && i > j) return;   // merged into the node its partner creates
    const uint2 f = flags[i];
    if (!mutual) {
        cout[f.x] = cin[i];
        return;
    }
    const uint32_t a = cin[i], b = cin[j];
    const uint32_t id = n + nodeBase + f.y;
    const float4 lo = min4(nlo[a], nlo[b]), hi = max4(nhi[a], nhi[b]);
    nlo[id] = lo;
    nhi[id] = hi;
    kids[id - n] = make_uint2(a, b);
    const uint32_t c = cnt[a] + cnt[b];
    const float A = box_area(lo, hi);
    const float split = 2.0f * A + cost[a] + cost[b];
    const float leafC = A * (float)c;
    const bool ok = c <= maxLeaf && leafC <= split;
    cnt[id] = c;
    cost[id] = ok ? leafC : split;
    leafOK[id] = ok ? 1u : 0u;
    cout[f.x] = id;
}

// Ends an iteration: the scan's totals give the surviving clusters and the new nodes; the node
// count after it goes to iterEnd[iteration].  No-op once one cluster is left or after a stall.
__global__ void k_ploc_advance(PlocState* st, const uint2* total, uint32_t* iterEnd, uint32_t iterCap) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    if (st->m <= 1 || st->stalled) return;
    const uint2 t = *total;
    if (t.y == 0 || t.x + t.y != st->m || st->iters >= iterCap) {
        st->stalled = 1u;
        return;
    }
    st->nodeCount += t.y;
    st->m = t.x;
    iterEnd[st->iters] = st->nodeCount;
    st->iters += 1u;
}

// ---- top-down: triangle slots (off = position of the subtree's first triangle), node states ----

__global__ void k_td_root(uint32_t root, const uint32_t* leafOK, uint32_t* state, uint32_t* off, uint32_t* depth) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    state[root] = leafOK[root] ? NODE_LEAF : NODE_KEPT;
    off[root] = 0u;
    depth[root] = 0u;
}

// Inner nodes n + [lo, hi): their states were set by their parents (created later).
__global__ __launch_bounds__(BB) void k_td(uint32_t n, uint32_t lo, uint32_t hi, const uint2* kids,
                                           const uint32_t* cnt, const uint32_t* leafOK, uint32_t* state,
                                           uint32_t* off, uint32_t* depth, uint32_t* bad) {
    const uint32_t k = lo + tid();
    if (k >= hi) return;
    const uint32_t X = n + k;
    const uint2 ch = kids[k];
    if (ch.x >= X || ch.y >= X) {   // a child is always created before its parent
        atomicOr(bad, 2u);
        return;
    }
    const uint32_t o = off[X], st = state[X], d = depth[X] + 1u;
    off[ch.x] = o;
    off[ch.y] = o + cnt[ch.x];
    depth[ch.x] = d;
    depth[ch.y] = d;
    if (st == NODE_KEPT) {
        state[ch.x] = leafOK[ch.x] ? NODE_LEAF : NODE_KEPT;
        state[ch.y] = leafOK[ch.y] ? NODE_LEAF : NODE_KEPT;
    } else {
        state[ch.x] = NODE_BELOW;
        state[ch.y] = NODE_BELOW;
    }
}

// Triangle of leaf s to its slot, 3 float4 per triangle (vertex 0's w = the primitive index,
// vertex 1's w = the skip bits, the complement of the triangle's CX_VIS_* mask); thread 0 also
// clears the TRI_PAD float4 after the last slot.
__global__ __launch_bounds__(BB) void k_tri_slots(const uint32_t* vals, const uint32_t* off, uint32_t n,
                                                  const float* pos, const uint32_t* idx, const uint8_t* mask,
                                                  float4* tris, uint32_t* bad) {
    const uint32_t s = tid();
    if (s >= n) return;
    if (s == 0)
        for (size_t k = 0; k < TRI_PAD; k++) tris[3 * (size_t)n + k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const uint32_t t = vals[s], slot = off[s];
    if (t >= n || slot >= n) {
        atomicOr(bad, 4u);
        return;
    }
    float4* o = tris + 3 * (size_t)slot;
    const uint32_t skip = mask ? (~(uint32_t)mask[t] & 3u) : 0u;
    for (int k = 0; k < 3; k++) {
        const float* v = pos + 3 * (size_t)idx[3 * (size_t)t + k];
        const uint32_t w = k == 0 ? t : (k == 1 ? skip : 0u);
        o[k] = make_float4(v[0], v[1], v[2], __builtin_bit_cast(float, w));
    }
}

__global__ __launch_bounds__(BB) void k_kept_flags(const uint32_t* state, uint32_t n, uint2* flags) {
    const uint32_t k = tid();
    if (k >= n - 1) return;
    flags[k] = make_uint2(state[n + k] == NODE_KEPT ? 1u : 0u, 0u);
}

__device__ __forceinline__ int32_t leaf_ref(uint32_t first, uint32_t count) {
    return (int32_t)~((first << 4) | count);
}

__device__ __forceinline__ int32_t child_ref(uint32_t c, uint32_t n, const uint32_t* state, const uint2* index,
                                             const uint32_t* off, const uint32_t* cnt) {
    if (state[c] == NODE_KEPT) return (int32_t)index[c - n].x;
    return leaf_ref(off[c], cnt[c]);
}

// Binary node per kept inner node (64 bytes, both children's padded boxes); the largest depth
// of a final leaf goes to *maxDepth.
__global__ __launch_bounds__(BB) void k_bin_nodes(uint32_t n, const uint2* kids, const uint32_t* state,
                                                  const uint2* index, const uint32_t* off, const uint32_t* cnt,
                                                  const uint32_t* depth, const float4* nlo, const float4* nhi,
                                                  float4* nodes, uint32_t numNodes, uint32_t* maxDepth,
                                                  uint32_t* bad) {
    __shared__ uint32_t s_d[BB / 64];
    const uint32_t k = tid();
    uint32_t dmax = 0u;
    bool mine = k < n - 1 && state[n + k] == NODE_KEPT;
    if (mine && index[k].x >= numNodes) {
        atomicOr(bad, 8u);
        mine = false;
    }
    if (mine) {
        const uint2 ch = kids[k];
        float4 lo[2] = {nlo[ch.x], nlo[ch.y]}, hi[2] = {nhi[ch.x], nhi[ch.y]};
        for (int c = 0; c < 2; c++) {
            float e = pad_of(lo[c].x, hi[c].x);
            lo[c].x -= e;
            hi[c].x += e;
            e = pad_of(lo[c].y, hi[c].y);
            lo[c].y -= e;
            hi[c].y += e;
            e = pad_of(lo[c].z, hi[c].z);
            lo[c].z -= e;
            hi[c].z += e;
        }
        float4* o = nodes + 4 * (size_t)index[k].x;
        o[0] = make_float4(lo[0].x, hi[0].x, lo[0].y, hi[0].y);
        o[1] = make_float4(lo[1].x, hi[1].x, lo[1].y, hi[1].y);
        o[2] = make_float4(lo[0].z, hi[0].z, lo[1].z, hi[1].z);
        const int32_t refs[4] = {child_ref(ch.x, n, state, index, off, cnt), child_ref(ch.y, n, state, index, off, cnt),
                                 0, 0};
        o[3] = __builtin_bit_cast(float4, refs);
        dmax = depth[n + k] + 1u;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dmax = max(dmax, (uint32_t)__shfl_xor(dmax, o, 64));
    if ((threadIdx.x & 63u) == 0) s_d[threadIdx.x >> 6] = dmax;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t r = s_d[0];
        for (int w = 1; w < BB / 64; w++) r = max(r, s_d[w]);
        if (r) atomicMax(maxDepth, r);
    }
}

// ---- wide collapse, one level of wide nodes per round ----

struct WChild {
    float lo[3], hi[3];
    int32_t ref;
};

__device__ __forceinline__ float wchild_area(const WChild& c) {
    const float dx = c.hi[0] - c.lo[0], dy = c.hi[1] - c.lo[1], dz = c.hi[2] - c.lo[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

__device__ __forceinline__ void bin_children(const float4* nodes, int32_t node, WChild& c0, WChild& c1) {
    const float4* q = nodes + 4 * (size_t)node;
    const float4 a = q[0], b = q[1], c = q[2];
    const int4 r = __builtin_bit_cast(int4, q[3]);
    c0 = WChild{{a.x, a.z, c.x}, {a.y, a.w, c.y}, r.x};
    c1 = WChild{{b.x, b.z, c.z}, {b.y, b.w, c.w}, r.y};
}

// Least-area collapse tables per binary node (the host build's dynamic programme, cx_bvh_build.cpp):
// cost[b][i] = least summed area of wide nodes for binary node b spread over at most i slots of
// its parent, use[b][i] = the slot count that optimum uses (1: b becomes a wide node), split[b][j]
// = the left child's share of j slots.  use == null selects the greedy rule.
struct WideDP {
    float* cost;
    uint8_t* use;
    uint8_t* split;
};

// Tables of the kept binary nodes made in one clustering iteration (ids n + [lo, hi)); their kept
// children were made in earlier iterations, so their tables are complete.
template <int N>
__global__ __launch_bounds__(BB) void k_wide_dp(uint32_t n, uint32_t lo, uint32_t hi, const uint32_t* state,
                                                const uint2* index, const float4* nodes, uint32_t numNodes,
                                                WideDP dp, uint32_t* bad) {
    const uint32_t k = lo + tid();
    if (k >= hi || state[n + k] != NODE_KEPT) return;
    const uint32_t b = index[k].x;
    if (b >= numNodes) {
        atomicOr(bad, 64u);
        return;
    }
    const float4* q = nodes + 4 * (size_t)b;
    const float4 x = q[0], y = q[1], z = q[2];
    const int4 r = __builtin_bit_cast(int4, q[3]);
    WChild c[2];
    c[0] = WChild{{x.x, x.z, z.x}, {x.y, x.w, z.y}, r.x};
    c[1] = WChild{{y.x, y.z, z.z}, {y.y, y.w, z.w}, r.y};
    if ((r.x >= 0 && (uint32_t)r.x >= b) || (r.y >= 0 && (uint32_t)r.y >= b)) {   // children come first
        atomicOr(bad, 64u);
        return;
    }
    WChild u;
    for (int a = 0; a < 3; a++) {
        u.lo[a] = fminf(c[0].lo[a], c[1].lo[a]);
        u.hi[a] = fmaxf(c[0].hi[a], c[1].hi[a]);
    }
    constexpr int S = N + 1;
    auto child_cost = [&](int ci, int i) -> float {
        return c[ci].ref < 0 ? 0.0f : dp.cost[(size_t)c[ci].ref * S + i];
    };
    float D[N + 1];
    const size_t base = (size_t)b * S;
    for (int j = 2; j <= N; j++) {
        D[j] = INFINITY;
        for (int kk = 1; kk < j; kk++) {
            const float v = child_cost(0, kk) + child_cost(1, j - kk);
            if (v < D[j]) {
                D[j] = v;
                dp.split[base + j] = (uint8_t)kk;
            }
        }
    }
    dp.cost[base + 1] = wchild_area(u) + D[N];
    dp.use[base + 1] = 1;
    for (int i = 2; i <= N; i++) {
        if (D[i] < dp.cost[base + i - 1]) {
            dp.cost[base + i] = D[i];
            dp.use[base + i] = (uint8_t)i;
        } else {
            dp.cost[base + i] = dp.cost[base + i - 1];
            dp.use[base + i] = dp.use[base + i - 1];
        }
    }
}

// Children of the wide node opened from binary node `node`.  With tables: the node spread over N
// slots by the least-area optimum.  Without: take its two children and keep replacing the inner
// child of largest surface area by its two children.
template <int N>
__device__ __forceinline__ int open_wide(const float4* nodes, uint32_t numNodes, int32_t node, WChild (&ch)[N],
                                        bool& ok, const WideDP& dp) {
    if (dp.use) {
        constexpr int S = N + 1;
        WChild pend[N + 1];
        int pendI[N + 1];
        int ps = 0, n = 0;
        auto distribute = [&](int32_t nd, int j) {
            WChild l, r;
            bin_children(nodes, nd, l, r);
            const int kk = dp.split[(size_t)nd * S + j];
            if (kk < 1 || kk >= j || ps + 2 > N + 1) {
                ok = false;
                return;
            }
            pend[ps] = r;
            pendI[ps++] = j - kk;
            pend[ps] = l;
            pendI[ps++] = kk;
        };
        distribute(node, N);
        while (ok && ps > 0) {
            const WChild c = pend[--ps];
            const int i = pendI[ps];
            if (c.ref < 0) {
                if (n < N) ch[n++] = c;
                else ok = false;
                continue;
            }
            if ((uint32_t)c.ref >= numNodes) {
                ok = false;
                break;
            }
            const int u = dp.use[(size_t)c.ref * S + i];
            if (u == 1) {
                if (n < N) ch[n++] = c;
                else ok = false;
            } else if (u >= 2 && u <= i) {
                distribute(c.ref, u);
            } else {
                ok = false;
            }
        }
        return n;
    }
    int n = 2;
    bin_children(nodes, node, ch[0], ch[1]);
    while (n < N) {
        int best = -1;
        float bestA = -1.0f;
        for (int i = 0; i < n; i++)
            if (ch[i].ref >= 0 && wchild_area(ch[i]) > bestA) {
                bestA = wchild_area(ch[i]);
                best = i;
            }
        if (best < 0) break;
        if ((uint32_t)ch[best].ref >= numNodes) {
            ok = false;
            break;
        }
        WChild a, b;
        bin_children(nodes, ch[best].ref, a, b);
        ch[best] = a;
        ch[n++] = b;
    }
    return n;
}

template <int N>
__global__ __launch_bounds__(BB) void k_wide_count(const uint2* front, uint32_t fN, const float4* nodes,
                                                   uint32_t numNodes, WideDP dp, uint2* counts, uint32_t* bad) {
    const uint32_t k = tid();
    if (k >= fN) return;
    if (front[k].x >= numNodes) {
        atomicOr(bad, 16u);
        counts[k] = make_uint2(0u, 0u);
        return;
    }
    WChild ch[N];
    bool ok = true;
    const int n = open_wide<N>(nodes, numNodes, (int32_t)front[k].x, ch, ok, dp);
    if (!ok) atomicOr(bad, 16u);
    uint32_t inner = 0u;
    for (int i = 0; i < n; i++) inner += ch[i].ref >= 0 ? 1u : 0u;
    counts[k] = make_uint2(inner, 0u);
}

// front[k] = (binary node, stack entries above it); counts[k].x = first slot of its inner
// children in the next level (after the scan).  Wide node k of the level is levelBase + k.
template <int N>
__global__ __launch_bounds__(BB) void k_wide_write(const uint2* front, uint32_t fN, uint32_t levelBase,
                                                   uint32_t nextBase, const uint2* counts, const float4* nodes,
                                                   uint32_t numNodes, WideDP dp, float4* wnodes, uint2* next,
                                                   uint32_t nextCap, uint32_t* maxStack, uint32_t* bad) {
    __shared__ uint32_t s_m[BB / 64];
    const uint32_t k = tid();
    uint32_t smax = 0u;
    bool mine = k < fN;
    // wide node levelBase + k must fit the array sized by numNodes (next-level slots are checked
    // against nextCap where they are written)
    if (mine && (front[k].x >= numNodes || levelBase + k >= numNodes)) {
        atomicOr(bad, 32u);
        mine = false;
    }
    if (mine) {
        WChild ch[N];
        const uint2 fk = front[k];
        bool ok = true;
        const int n = open_wide<N>(nodes, numNodes, (int32_t)fk.x, ch, ok, dp);
        if (!ok) atomicOr(bad, 32u);
        const uint32_t stack = fk.y + (uint32_t)(n - 1);
        smax = stack;
        uint32_t slot = counts[k].x;
        int32_t refs[N];
        for (int i = 0; i < N; i++) {
            refs[i] = WIDE_EMPTY;
            if (i >= n) continue;
            if (ch[i].ref >= 0) {
                if (slot < nextCap) next[slot] = make_uint2((uint32_t)ch[i].ref, stack);
                else atomicOr(bad, 32u);
                refs[i] = (int32_t)(nextBase + slot);
                slot++;
            } else {
                refs[i] = ch[i].ref;
            }
        }
        float4* w = wnodes + (size_t)(levelBase + k) * (2 * N);
        for (int g = 0; g < N / 4; g++) {
            float v[6][4];
            int32_t r4[4];
            for (int j = 0; j < 4; j++) {
                const int i = 4 * g + j;
                for (int a = 0; a < 3; a++) {
                    v[2 * a][j] = i < n ? ch[i].lo[a] : WIDE_EMPTY_COORD;
                    v[2 * a + 1][j] = i < n ? ch[i].hi[a] : WIDE_EMPTY_COORD;
                }
                r4[j] = refs[i];
            }
            for (int q = 0; q < 6; q++) w[8 * g + q] = make_float4(v[q][0], v[q][1], v[q][2], v[q][3]);
            w[8 * g + 6] = __builtin_bit_cast(float4, r4);
            w[8 * g + 7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) smax = max(smax, (uint32_t)__shfl_xor(smax, o, 64));
    if ((threadIdx.x & 63u) == 0) s_m[threadIdx.x >> 6] = smax;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t r = s_m[0];
        for (int w = 1; w < BB / 64; w++) r = max(r, s_m[w]);
        if (r) atomicMax(maxStack, r);
    }
}

// ---- validation: every reference empty, an inner node in range, or a leaf inside the slots ----

__device__ __forceinline__ bool ref_ok(int32_t r, uint32_t numNodes, uint32_t slots) {
    if (r == WIDE_EMPTY) return true;
    if (r >= 0) return (uint32_t)r < numNodes;
    const uint32_t code = ~(uint32_t)r;
    return (code & 15u) != 0u && (uint64_t)(code >> 4) + (code & 15u) <= slots;
}

__global__ __launch_bounds__(BB) void k_check_refs(const float4* nodes, uint32_t numNodes, int words, uint32_t slots,
                                                   uint32_t* bad) {
    const uint32_t k = tid();
    if (k >= numNodes) return;
    bool ok = true;
    if (words == 4) {
        const int4 r = __builtin_bit_cast(int4, nodes[4 * (size_t)k + 3]);
        ok = r.x != WIDE_EMPTY && r.y != WIDE_EMPTY && ref_ok(r.x, numNodes, slots) && ref_ok(r.y, numNodes, slots);
    } else {
        for (int g = 0; g < words / 8; g++) {
            const int4 r = __builtin_bit_cast(int4, nodes[(size_t)k * words + 8 * g + 6]);
            ok = ok && ref_ok(r.x, numNodes, slots) && ref_ok(r.y, numNodes, slots) && ref_ok(r.z, numNodes, slots) &&
                 ref_ok(r.w, numNodes, slots);
        }
    }
    if (!ok) atomicOr(bad, 1u);
}

// Carves device buffers out of one allocation, 256-byte aligned.
struct Arena {
    char* base = nullptr;
    size_t used = 0;
    template <typename T>
    T* take(size_t count) {
        T* p = (T*)(base + used);
        used += (count * sizeof(T) + 255) & ~(size_t)255;
        return p;
    }
};

struct Scan {
    uint2* tileSums;
    uint2* total;
    // n elements, or *nDev elements (at most n) when nDev is given
    hipError_t run(uint2* a, uint32_t n, hipStream_t s, const uint32_t* nDev = nullptr) const {
        if (n == 0) return hipMemsetAsync(total, 0, sizeof(uint2), s);
        const uint32_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
        hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(BB), 0, s, a, n, nDev, tileSums);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(BB), 0, s, tileSums, tiles, nDev, total);
        hipLaunchKernelGGL(k_scan_apply, dim3(tiles), dim3(BB), 0, s, a, n, nDev, tileSums);
        return hipGetLastError();
    }
};

inline dim3 blocks(uint32_t n) { return dim3((std::max(n, 1u) + BB - 1) / BB); }

}  // namespace

bool build_bvh_device(const float* pos, const uint32_t* idx, const uint8_t* mask, uint32_t n, uint32_t maxLeaf,
                      int wantWidth, DeviceBvh& out, std::vector<void*>& allocs, std::string& err, hipError_t& herr) {
    out = DeviceBvh{};
    herr = hipSuccess;
    if (n == 0) return true;
    if (n >= (1u << 27)) {
        err = "scene has too many triangles for the leaf encoding (limit 2^27)";
        return false;
    }
    maxLeaf = std::max(1u, std::min(15u, maxLeaf));
    const hipStream_t s = nullptr;
    const uint32_t nodesAll = 2 * n - 1, inner = n - 1;
    const uint32_t triBlocks = blocks(n).x;
    const size_t tilesMax = (n + SCAN_TILE - 1) / SCAN_TILE;
    // one temporary allocation for every build buffer
    Arena ar;
    auto plan = [&](Arena& A) {
        struct {
            float4 *triLo, *triHi, *nlo, *nhi;
            uint32_t *keys, *vals, *keysAlt, *valsAlt, *numValid, *cA, *cB, *nn, *cnt, *leafOK, *state, *off, *depth;
            uint32_t *scalars, *iterEnd;
            PlocState* ploc;
            float *partial, *bounds, *cost;
            uint2 *kids, *flags, *tileSums, *total, *frontA, *frontB;
            void* sortTemp;
        } b;
        b.triLo = A.take<float4>(n);
        b.triHi = A.take<float4>(n);
        b.nlo = A.take<float4>(nodesAll);
        b.nhi = A.take<float4>(nodesAll);
        b.keys = A.take<uint32_t>(n);
        b.vals = A.take<uint32_t>(n);
        b.keysAlt = A.take<uint32_t>(n);
        b.valsAlt = A.take<uint32_t>(n);
        b.numValid = A.take<uint32_t>(1);
        b.cA = A.take<uint32_t>(n);
        b.cB = A.take<uint32_t>(n);
        b.nn = A.take<uint32_t>(n);
        b.cnt = A.take<uint32_t>(nodesAll);
        b.leafOK = A.take<uint32_t>(nodesAll);
        b.state = A.take<uint32_t>(nodesAll);
        b.off = A.take<uint32_t>(nodesAll);
        b.depth = A.take<uint32_t>(nodesAll);
        b.scalars = A.take<uint32_t>(4);
        b.ploc = A.take<PlocState>(1);
        b.iterEnd = A.take<uint32_t>(n);
        b.partial = A.take<float>(6 * (size_t)triBlocks);
        b.bounds = A.take<float>(6);
        b.cost = A.take<float>(nodesAll);
        b.kids = A.take<uint2>(std::max(inner, 1u));
        b.flags = A.take<uint2>(n);
        b.tileSums = A.take<uint2>(tilesMax);
        b.total = A.take<uint2>(1);
        b.frontA = A.take<uint2>(n);
        b.frontB = A.take<uint2>(n);
        b.sortTemp = A.take<char>(sort_temp_bytes(n));
        return b;
    };
    {
        Arena sizing;
        (void)plan(sizing);
        herr = hipMalloc((void**)&ar.base, sizing.used);
        if (herr != hipSuccess) {
            err = "device BVH build: allocation of the build buffers failed";
            return false;
        }
    }
    auto B = plan(ar);
    struct Guard {   // frees the build buffers once the stream has finished with them
        void* p;
        ~Guard() {
            (void)hipStreamSynchronize(nullptr);
            (void)hipFree(p);
        }
    } guard{ar.base};
    auto fail = [&](hipError_t e, const char* what) {
        herr = e;
        err = std::string("device BVH build: ") + what;
        return false;
    };
    hipError_t e = hipMemsetAsync(B.scalars, 0, 4 * sizeof(uint32_t), s);
    if (e != hipSuccess) return fail(e, "memset");
    // boxes, Morton keys, sort
    hipLaunchKernelGGL(k_tri_boxes, dim3(triBlocks), dim3(BB), 0, s, pos, idx, n, B.triLo, B.triHi, B.partial);
    hipLaunchKernelGGL(k_bounds, dim3(1), dim3(BB), 0, s, B.partial, triBlocks, B.bounds);
    hipLaunchKernelGGL(k_morton, blocks(n), dim3(BB), 0, s, B.triLo, B.triHi, n, B.bounds, B.keys);
    e = launch_key_sort(B.keys, B.vals, B.keysAlt, B.valsAlt, n, B.numValid, B.sortTemp, sort_temp_bytes(n), s);
    if (e != hipSuccess) return fail(e, "sort");
    const uint32_t* order = B.vals;   // leaf s holds triangle order[s]
    hipLaunchKernelGGL(k_ploc_init, blocks(n), dim3(BB), 0, s, order, B.triLo, B.triHi, n, B.nlo, B.nhi, B.cnt,
                       B.cost, B.leafOK, B.cA);
    // Clustering.  The cluster count lives on the device (PlocState); the host enqueues
    // PLOC_BATCH iterations at a time, sized for the last count it read (an upper bound: the
    // count only falls), and reads the state back once per batch.
    const Scan scan{B.tileSums, B.total};
    const PlocState st0{n, 0u, 0u, 0u};
    e = hipMemcpyAsync(B.ploc, &st0, sizeof(st0), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return fail(e, "clustering state");
    PlocState st = st0;
    uint32_t* cin = B.cA;
    uint32_t* cout = B.cB;
    constexpr int PLOC_BATCH = 16;
    while (st.m > 1) {
        for (int it = 0; it < PLOC_BATCH && e == hipSuccess; it++) {
            hipLaunchKernelGGL(k_ploc_nn, blocks(st.m), dim3(BB), 0, s, cin, B.ploc, B.nlo, B.nhi, B.nn);
            hipLaunchKernelGGL(k_ploc_flags, blocks(st.m), dim3(BB), 0, s, B.nn, B.ploc, B.flags);
            e = scan.run(B.flags, st.m, s, &B.ploc->m);
            hipLaunchKernelGGL(k_ploc_merge, blocks(st.m), dim3(BB), 0, s, cin, cout, B.nn, B.flags, B.ploc, n, maxLeaf,
                               B.nlo, B.nhi, B.kids, B.cnt, B.cost, B.leafOK);
            hipLaunchKernelGGL(k_ploc_advance, dim3(1), dim3(64), 0, s, B.ploc, B.total, B.iterEnd, n);
            std::swap(cin, cout);
        }
        if (e == hipSuccess) e = hipMemcpy(&st, B.ploc, sizeof(st), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return fail(e, "clustering");
        if (st.stalled) {
            herr = hipErrorUnknown;
            err = "device BVH build did not converge";
            return false;
        }
    }
    if (st.nodeCount != inner) return fail(hipErrorUnknown, "clustering left the wrong number of inner nodes");
    std::vector<uint32_t> iterEnd(st.iters);   // inner nodes created after each iteration
    if (st.iters) e = hipMemcpy(iterEnd.data(), B.iterEnd, st.iters * sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(e, "clustering");
    const uint32_t root = n == 1 ? 0u : n + inner - 1;   // the last node created
    // top-down: slots, states, depths
    hipLaunchKernelGGL(k_td_root, dim3(1), dim3(64), 0, s, root, B.leafOK, B.state, B.off, B.depth);
    for (size_t it = iterEnd.size(); it-- > 0;) {
        const uint32_t lo = it ? iterEnd[it - 1] : 0u, hi = iterEnd[it];
        hipLaunchKernelGGL(k_td, blocks(hi - lo), dim3(BB), 0, s, n, lo, hi, B.kids, B.cnt, B.leafOK, B.state, B.off,
                           B.depth, B.scalars + 2);
    }
    // kept inner nodes in creation order
    uint32_t numNodes = 0;
    if (inner) {
        hipLaunchKernelGGL(k_kept_flags, blocks(inner), dim3(BB), 0, s, B.state, n, B.flags);
        e = scan.run(B.flags, inner, s);
        if (e != hipSuccess) return fail(e, "scan");
        uint2 tot;
        e = hipMemcpy(&tot, B.total, sizeof(tot), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return fail(e, "node count");
        numNodes = tot.x;
    }
    auto alloc = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (herr != hipSuccess) return nullptr;
        herr = hipMalloc(&p, std::max<size_t>(bytes, 16));
        if (herr != hipSuccess) return nullptr;
        allocs.push_back(p);
        return p;
    };
    out.tris = (float4*)alloc((3 * (size_t)n + TRI_PAD) * sizeof(float4));
    out.nodes = (float4*)alloc(4 * (size_t)numNodes * sizeof(float4));
    if (herr != hipSuccess) return fail(herr, "allocation of the node arrays failed");
    uint32_t* maxDepth = B.scalars;
    uint32_t* maxStack = B.scalars + 1;
    uint32_t* bad = B.scalars + 2;
    hipLaunchKernelGGL(k_tri_slots, blocks(n), dim3(BB), 0, s, order, B.off, n, pos, idx, mask, out.tris, bad);
    if (numNodes)
        hipLaunchKernelGGL(k_bin_nodes, blocks(inner), dim3(BB), 0, s, n, B.kids, B.state, B.flags, B.off, B.cnt,
                           B.depth, B.nlo, B.nhi, out.nodes, numNodes, maxDepth, bad);
    out.numNodes = numNodes;
    // the root is the last kept node (the largest id), or a single leaf over every triangle
    out.root = numNodes ? (int32_t)(numNodes - 1) : (int32_t)~((0u << 4) | n);
    out.wroot = out.root;
    out.width = 2;
    uint32_t sc[4];
    e = hipMemcpy(sc, B.scalars, sizeof(sc), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(e, "depth");
    if (sc[2]) return fail(hipErrorUnknown, "inconsistent tree (internal error)");
    out.maxDepth = sc[0];
    if (out.maxDepth + 1 >= (uint32_t)STACK_DEPTH) {
        herr = hipErrorUnknown;
        out.tooDeep = true;
        err = "device BVH deeper than the traversal stack";
        return false;
    }
    out.stackBound = out.maxDepth;
    if ((wantWidth == 4 || wantWidth == 8) && numNodes == 0) {
        out.width = (uint32_t)wantWidth;   // a single leaf: no inner node at any width
    } else if (wantWidth == 4 || wantWidth == 8) {
        const int words = 2 * wantWidth;
        const size_t S = (size_t)wantWidth + 1;
        float4* wnodes = nullptr;
        WideDP dp{nullptr, nullptr, nullptr};
        e = hipMalloc((void**)&wnodes, (size_t)numNodes * words * sizeof(float4));
        if (e == hipSuccess) e = hipMalloc((void**)&dp.cost, (size_t)numNodes * S * sizeof(float));
        if (e == hipSuccess) e = hipMalloc((void**)&dp.use, (size_t)numNodes * S);
        if (e == hipSuccess) e = hipMalloc((void**)&dp.split, (size_t)numNodes * S);
        struct Tables {   // the programme's tables live only through the collapse
            WideDP& t;
            ~Tables() {
                (void)hipStreamSynchronize(nullptr);
                (void)hipFree(t.cost);
                (void)hipFree(t.use);
                (void)hipFree(t.split);
            }
        } tables{dp};
        if (e != hipSuccess) {
            (void)hipFree(wnodes);
            return fail(e, "allocation of the wide nodes failed");
        }
        // the tables, one clustering iteration at a time (B.flags still holds the kept ranks)
        for (size_t it = 0; it < iterEnd.size(); it++) {
            const uint32_t lo = it ? iterEnd[it - 1] : 0u, hi = iterEnd[it];
            if (wantWidth == 4)
                hipLaunchKernelGGL(k_wide_dp<4>, blocks(hi - lo), dim3(BB), 0, s, n, lo, hi, B.state, B.flags, out.nodes,
                                   numNodes, dp, bad);
            else
                hipLaunchKernelGGL(k_wide_dp<8>, blocks(hi - lo), dim3(BB), 0, s, n, lo, hi, B.state, B.flags, out.nodes,
                                   numNodes, dp, bad);
        }
        // one level of wide nodes per round; the least-area tables first, the greedy rule if
        // their tree's stack bound does not fit
        uint32_t numWide = 0;
        auto collapse = [&](const WideDP& rule) -> hipError_t {
            hipError_t ce = hipMemsetAsync(maxStack, 0, sizeof(uint32_t), s);
            uint2 rootEntry = make_uint2((uint32_t)out.root, 0u);
            if (ce == hipSuccess) ce = hipMemcpy(B.frontA, &rootEntry, sizeof(rootEntry), hipMemcpyHostToDevice);
            uint2* front = B.frontA;
            uint2* next = B.frontB;
            uint32_t fN = 1, levelBase = 0;
            while (fN && ce == hipSuccess) {
                if (wantWidth == 4)
                    hipLaunchKernelGGL(k_wide_count<4>, blocks(fN), dim3(BB), 0, s, front, fN, out.nodes, numNodes, rule,
                                       B.flags, bad);
                else
                    hipLaunchKernelGGL(k_wide_count<8>, blocks(fN), dim3(BB), 0, s, front, fN, out.nodes, numNodes, rule,
                                       B.flags, bad);
                ce = scan.run(B.flags, fN, s);
                if (ce != hipSuccess) break;
                const uint32_t nextBase = levelBase + fN;
                if (wantWidth == 4)
                    hipLaunchKernelGGL(k_wide_write<4>, blocks(fN), dim3(BB), 0, s, front, fN, levelBase, nextBase,
                                       B.flags, out.nodes, numNodes, rule, wnodes, next, n, maxStack, bad);
                else
                    hipLaunchKernelGGL(k_wide_write<8>, blocks(fN), dim3(BB), 0, s, front, fN, levelBase, nextBase,
                                       B.flags, out.nodes, numNodes, rule, wnodes, next, n, maxStack, bad);
                uint2 tot;
                ce = hipMemcpy(&tot, B.total, sizeof(tot), hipMemcpyDeviceToHost);
                if (ce != hipSuccess) break;
                if (nextBase + tot.x > numNodes) {
                    ce = hipErrorUnknown;
                    break;
                }
                levelBase = nextBase;
                fN = tot.x;
                std::swap(front, next);
            }
            numWide = levelBase;
            if (ce == hipSuccess) ce = hipMemcpy(sc, B.scalars, sizeof(sc), hipMemcpyDeviceToHost);
            return ce;
        };
        e = collapse(dp);
        if (e == hipSuccess && sc[1] + 1 > (uint32_t)STACK_DEPTH) e = collapse(WideDP{nullptr, nullptr, nullptr});
        if (e == hipSuccess && sc[1] + 1 <= (uint32_t)STACK_DEPTH) {
            hipLaunchKernelGGL(k_check_refs, blocks(numWide), dim3(BB), 0, s, wnodes, numWide, words, n, bad);
            allocs.push_back(wnodes);
            out.wnodes = wnodes;
            out.numWide = numWide;
            out.wroot = 0;
            out.width = (uint32_t)wantWidth;
            out.stackBound = sc[1];
        } else {
            (void)hipFree(wnodes);
            if (e != hipSuccess) return fail(e, "wide collapse");
        }
    }
    if (numNodes) hipLaunchKernelGGL(k_check_refs, blocks(numNodes), dim3(BB), 0, s, out.nodes, numNodes, 4, n, bad);
    e = hipMemcpy(sc, B.scalars, sizeof(sc), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(e, "validation");
    if (sc[2]) return fail(hipErrorUnknown, "a node reference is out of range");
    return true;
}

}  // namespace cx
