// Internal declarations shared by the kernel translation unit and the host library code.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/crisoptix.h"
#include "cx_restir_dev.h"
#include "cx_wave.h"

namespace cx {

hipError_t launch_pass(int pass, const FrameArgs& A, hipStream_t s);
hipError_t launch_reset(const FrameArgs& A, hipStream_t s);
hipError_t launch_zero_counters(unsigned long long* c, uint32_t n, hipStream_t s);
hipError_t launch_set_dyn(FrameDyn* dst, const FrameDyn& d, hipStream_t s);
hipError_t launch_next_dyn(FrameDyn* dst, uint64_t* state, const FrameDyn& d, hipStream_t s);
hipError_t launch_set_materials(DMaterial* dst, const float* rows, uint32_t n, hipStream_t s);
hipError_t launch_store_entries(const FrameArgs& A, const EntryBuild& E, hipStream_t s);
hipError_t launch_entries_from_arrays(const float* pos, const float* r, const uint8_t* valid, uint32_t n,
                                      const EntryBuild& E, hipStream_t s);
size_t sort_temp_bytes(uint32_t n);
hipError_t launch_key_sort(uint32_t* keys, uint32_t* vals, uint32_t* keysAlt, uint32_t* valsAlt, uint32_t n,
                           uint32_t* numValid, void* temp, size_t tempBytes, hipStream_t s);
hipError_t launch_tree_build(const EntryBuild& E, const EntryTree& T, uint32_t n, void* temp, size_t tempBytes,
                             hipStream_t s);
hipError_t launch_entries_query(const EntryTree& T, const float* q, uint32_t nq, uint32_t k, uint32_t* idx,
                                float* dist, uint32_t* count, hipStream_t s);
// Morton quantisation of the entry positions over the box [mn, mx] (10 bits per axis).
inline void set_quantisation(EntryBuild& E, const float* mn, const float* mx) {
    E.lo = V3{mn[0], mn[1], mn[2]};
    float sc[3];
    for (int a = 0; a < 3; a++) {
        const float w = mx[a] - mn[a];
        sc[a] = w > 0.0f ? 1023.0f / w : 0.0f;
    }
    E.scale = V3{sc[0], sc[1], sc[2]};
}

// Static scratch of cx_entries_search_static (cx_noalloc.hip): takes the address of the
// file-scope device array on the current device, once per device.  Called by cx_ctx_create.
hipError_t noalloc_setup(int device);
// The context's last-error message and device, for entry points outside cx_host.cpp.
void ctx_set_error(cx_ctx* ctx, const char* msg);
int ctx_device(const cx_ctx* ctx);

hipError_t launch_trace(const DScene& S, const float* rays, uint32_t n, int anyHit, uint32_t* prim, float* t,
                        float* u, float* v, uint32_t* occ, hipStream_t s);

// Path-replay adjoint (cx_adjoint.hip).  Gradients accumulate into the double arrays.
struct AdjArgs {
    const float* dLdI;             // [P][3] dL/d(pixel estimate), or null (replay only)
    double* gAlbedo;               // [materials][3] or null
    double* gRough;                // [materials] or null
    double* gEmit;                 // [emitters][3] or null
    float* replay;                 // [P][3] the replayed estimate, or null
    uint32_t* status;              // [3] replayed, replay differs from the record, skipped
    const DMaterial* evalMats;     // values and derivatives use these materials
    const float* evalEmit;         // and these emitter colours
    uint32_t numMaterials, numEmitters;
    // Gradient slots summed per block in LDS before one global add each (0: every term is a
    // global add): albedo [0, 3M), roughness [3M, 4M), emitter colour [4M, 4M + 3E).
    uint32_t ldsSlots;
};
hipError_t launch_adjoint(const FrameArgs& A, const AdjArgs& G, hipStream_t s);
hipError_t launch_tonemap(const float* hdr, uint8_t* ldr, uint32_t n, hipStream_t s);

// Wavefront kernels (cx_wave.hip).  Step kernels (*_STEP) run grid-stride over the jobs of
// W.qin; the others run one thread per pixel (threads = P).
enum WaveKernel {
    WK_PREFIX_INIT,
    WK_PG_STEP,
    WK_CP_STEP,
    WK_PREFIX_TEMPORAL_SETUP,
    WK_SH_STEP,
    WK_PREFIX_RESAMPLE,
    WK_PREFIX_SPATIAL_SETUP,
    WK_PREFIX_SPATIAL_RESAMPLE,
    WK_SUFFIX_INIT,
    WK_GATHER_SUFFIX_INIT,
    WK_SG_STEP,
    WK_SUFFIX_TEMPORAL_SETUP,
    WK_SUFFIX_RESAMPLE,
    WK_SUFFIX_SPATIAL_SETUP,
    WK_SUFFIX_SPATIAL_RESAMPLE,
    WK_GATHER_INIT,
    WK_GATHER_SETUP,
    WK_GATHER_COMBINE,
    WK_GATHER_COMBINE_LAST,
    WK_FG_INIT,
    WK_FG_STEP,
    WK_SUFFIX_AOS,
};

hipError_t wave_zero_counts(uint32_t* c, hipStream_t s);
hipError_t wave_tail(int step, const WaveArgs& W, uint32_t maxSteps, int pass, hipStream_t s);
// minChunk: the fewest rays per traversal chunk (1 for a frame that has the device alone, the
// largest chunk for one that runs beside others; results are the same either way).
hipError_t wave_trace(const DScene& S, const Queue& qin, const Queue& qout, unsigned long long* counters,
                      uint32_t minChunk, hipStream_t s);
hipError_t wave_launch(int kernel, const WaveArgs& W, uint32_t threads, hipStream_t s);

// Most triangles per BVH leaf when cx_scene_desc.bvhMaxLeafSize is 0.  Measured on the
// 1.06 M-triangle object grid (4-wide, 1080p Conditional ReSTIR): leaves of at most 1 / 2 / 4 /
// 8 triangles trace a frame's rays in 5.13-5.18 / 4.87-4.88 / 5.03-5.05 / 5.15 ms.
static constexpr uint32_t DEFAULT_MAX_LEAF = 2;

// Host BVH build output (binary SAH BVH in the node format of cx_scene_dev.h).
// Zero float4 after the last triangle slot of every triangle array the traversal reads: the
// 4-wide traversal step loads 7 float4 from a leaf's first slot whatever the leaf's size, so the
// loads of a last leaf of one triangle end 4 float4 past the slots.
static constexpr size_t TRI_PAD = 4;

struct BvhHost {
    std::vector<float4> nodes;   // 4 per inner node
    std::vector<float4> tris;    // 3 per triangle slot, then TRI_PAD zero float4 (none without triangles)
    int32_t root = 0;
    uint32_t maxDepth = 0;
};

// mask: CX_VIS_* bits per triangle (host), or null; stored as skip bits in the triangle slots.
bool build_bvh(const float* pos, const uint32_t* idx, uint32_t numTris, uint32_t maxLeaf, BvhHost& out,
               std::string& err, const uint8_t* mask = nullptr);

// Device BVH build (cx_bvh_device.hip) from device positions and indices: binary nodes, triangle
// slots and, when wantWidth is 4 or 8 and its stack bound fits, wide nodes, in the formats of
// the host build.  The arrays it returns are appended to `allocs` (owned by the caller).
struct DeviceBvh {
    float4* nodes = nullptr;     // 4 float4 per binary inner node
    float4* wnodes = nullptr;    // 2 * width float4 per wide node, or null
    float4* tris = nullptr;      // 3 float4 per triangle slot
    uint32_t numNodes = 0, numWide = 0;
    int32_t root = 0, wroot = 0;
    uint32_t width = 2, maxDepth = 0, stackBound = 0;
    bool tooDeep = false;        // the binary tree is deeper than the traversal stack
};
// mask: CX_VIS_* bits per triangle in device memory, or null.
bool build_bvh_device(const float* pos, const uint32_t* idx, const uint8_t* mask, uint32_t numTris, uint32_t maxLeaf,
                      int wantWidth, DeviceBvh& out, std::vector<void*>& allocs, std::string& err, hipError_t& herr);

// Wide (4- or 8-child) node array from a binary BVH; false if its worst-case traversal stack
// would not fit in `stackLimit` entries.
bool collapse_bvh(const BvhHost& bin, int width, uint32_t stackLimit, std::vector<float4>& out, int32_t& root,
                  uint32_t& stackBound);

}  // namespace cx
