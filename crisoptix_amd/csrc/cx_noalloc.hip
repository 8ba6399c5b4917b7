// cx_entries_search_static: the prefix-entry structure built and queried in one call with no
// device allocation (include/crisoptix.h).  It runs the kernels of cx_entries_build followed by
// cx_entries_query (entry records, the memset-free radix sort, the linear BVH over the sorted
// entries with its range-tree bounds, the k-nearest query); the buffers those keep in a
// cx_entries object live here in one file-scope device array of fixed size instead.
//
// The array's name is a host-side stand-in in host code, not a device address: its address on
// each device is taken once with hipGetSymbolAddress (noalloc_setup, called by cx_ctx_create
// with the device current) and only that address is passed to kernels.  Every call on a device
// shares that device's array and leaves its contents behind; calls that overlap on one device
// (two streams, say) race on it.
#include <mutex>

#include "cx_internal.h"

namespace cx {
namespace {

constexpr uint32_t NOALLOC_MAX_ENTRIES = 65536;   // a 256 x 256 frame's prefix entries
constexpr uint32_t NOALLOC_MAX_DEVICES = 64;

constexpr size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Byte offsets of the buffers inside the array.
struct Layout {
    size_t pos, valid, keys, vals, keysAlt, valsAlt, numValid;
    size_t entrySorted, entryPixel, nodes, segLo, segHi, segLevels, sortTemp, sortTempBytes, total;
};

constexpr Layout make_layout() {
    constexpr size_t n = NOALLOC_MAX_ENTRIES;
    // the sort's per-digit tile counts and digit totals (sort_temp_bytes: tiles of 1024 keys)
    constexpr size_t sortBytes = (256 * ((n + 1023) / 1024) + 256) * sizeof(uint32_t);
    Layout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) {
        const size_t at = o;
        o = align256(o + bytes);
        return at;
    };
    L.pos = take(n * sizeof(float4));
    L.valid = take(n);
    L.keys = take(n * sizeof(uint32_t));
    L.vals = take(n * sizeof(uint32_t));
    L.keysAlt = take(n * sizeof(uint32_t));
    L.valsAlt = take(n * sizeof(uint32_t));
    L.numValid = take(sizeof(uint32_t));
    L.entrySorted = take(n * sizeof(float4));
    L.entryPixel = take(n * sizeof(uint32_t));
    L.nodes = take(4 * n * sizeof(float4));
    L.segLo = take(((size_t(2) << entry_seg_levels(n)) - 1) * sizeof(float4));
    L.segHi = take(((size_t(2) << entry_seg_levels(n)) - 1) * sizeof(float4));
    L.segLevels = take(sizeof(uint32_t));
    L.sortTemp = take(sortBytes);
    L.sortTempBytes = sortBytes;
    L.total = o;
    return L;
}

constexpr Layout LAYOUT = make_layout();
static_assert(LAYOUT.total <= (size_t(9) << 20), "static scratch stays below 9 MiB");

std::mutex g_mu;
void* g_ptr[NOALLOC_MAX_DEVICES] = {};   // the array's address on each device, once taken

}  // namespace
}  // namespace cx

static __device__ __align__(256) unsigned char scratch[cx::LAYOUT.total];

namespace cx {

hipError_t noalloc_setup(int device) {
    if (device < 0 || device >= (int)NOALLOC_MAX_DEVICES) return hipSuccess;   // the entry point refuses it
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_ptr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[device] = p;
    return hipSuccess;
}

}  // namespace cx

using namespace cx;

extern "C" int cx_entries_search_static(cx_ctx* ctx, const float* pos, const float* radius, const uint8_t* valid,
                                        uint32_t n, const float* sceneMin, const float* sceneMax, const float* query,
                                        uint32_t nq, uint32_t k, uint32_t* idx, float* dist, uint32_t* count,
                                        void* stream) {
    if (!ctx) return CX_ERR_INVALID_ARG;
    auto fail = [&](int code, const char* msg) {
        ctx_set_error(ctx, msg);
        return code;
    };
    if (!sceneMin || !sceneMax || (n && (!pos || !radius || !valid)) || (nq && (!query || !idx || !dist || !count)))
        return fail(CX_ERR_INVALID_ARG, "null argument");
    if (n > NOALLOC_MAX_ENTRIES)
        return fail(CX_ERR_INVALID_ARG, "cx_entries_search_static: more than 65536 entries (the static scratch's "
                                        "capacity); use cx_entries_create/cx_entries_build for larger sets");
    if (k == 0 || k > 31) return fail(CX_ERR_INVALID_ARG, "k must be in [1, 31]");
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev != ctx_device(ctx))
        return fail(CX_ERR_INVALID_ARG, "cx_entries_search_static: the context's device is not the current device");
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        if (dev >= 0 && dev < (int)NOALLOC_MAX_DEVICES) base = g_ptr[dev];
    }
    if (!base) return fail(CX_ERR_UNSUPPORTED, "cx_entries_search_static: no static scratch on this device");
    if (sort_temp_bytes(n) > LAYOUT.sortTempBytes) return fail(CX_ERR_INVALID_ARG, "sort scratch too small");
    unsigned char* b = (unsigned char*)base;
    EntryBuild E{};
    E.pos = (float4*)(b + LAYOUT.pos);
    E.valid = (uint8_t*)(b + LAYOUT.valid);
    E.keys = (uint32_t*)(b + LAYOUT.keys);
    E.vals = (uint32_t*)(b + LAYOUT.vals);
    E.keysAlt = (uint32_t*)(b + LAYOUT.keysAlt);
    E.valsAlt = (uint32_t*)(b + LAYOUT.valsAlt);
    E.numValid = (uint32_t*)(b + LAYOUT.numValid);
    set_quantisation(E, sceneMin, sceneMax);
    EntryTree T{};
    T.entrySorted = (float4*)(b + LAYOUT.entrySorted);
    T.entryPixel = (uint32_t*)(b + LAYOUT.entryPixel);
    T.nodes = (float4*)(b + LAYOUT.nodes);
    T.segLo = (float4*)(b + LAYOUT.segLo);
    T.segHi = (float4*)(b + LAYOUT.segHi);
    T.segLevels = (uint32_t*)(b + LAYOUT.segLevels);
    T.segLevelsCap = entry_seg_levels(n > 0 ? n : 1u);   // <= the layout's levels: the arrays fit
    T.sortedKeys = E.keys;   // the sort leaves its output in keys / vals
    T.numValid = E.numValid;
    const hipStream_t s = (hipStream_t)stream;
    // every value the kernels read is written earlier in this call: the entry records for all
    // n slots, the valid count (cleared by a fill kernel, then set by the sort's first pass), the
    // sorted slots, the range tree's levels and nodes over the valid entries, and the LBVH nodes
    hipError_t e = launch_entries_from_arrays(pos, radius, valid, n, E, s);
    if (e == hipSuccess) e = launch_tree_build(E, T, n, b + LAYOUT.sortTemp, LAYOUT.sortTempBytes, s);
    if (e == hipSuccess) e = launch_entries_query(T, query, nq, k, idx, dist, count, s);
    if (e != hipSuccess) {
        ctx_set_error(ctx, (std::string("cx_entries_search_static: ") + hipGetErrorString(e)).c_str());
        return CX_ERR_HIP;
    }
    return CX_OK;
}
