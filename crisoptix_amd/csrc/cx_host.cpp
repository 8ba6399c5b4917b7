// C ABI of the frame loop (include/crisoptix.h): context, scene upload + BVH, the prefix-entry
// search structure, and the renderer that runs Renderer::LaunchFrame's phase order
// (src/Renderer.cpp:209-406) as HIP kernels on the caller's stream with no device sync.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "cx_internal.h"

using namespace cx;

struct cx_ctx {
    int device = 0;
    std::string err;
    std::vector<hipStream_t> pool;   // streams of cx_batch_frame, created on first use
};

struct cx_scene {
    cx_ctx* ctx = nullptr;
    DScene d{};
    DMaterial* evalMats = nullptr;   // evaluation table of the adjoint (device)
    float* evalEmit = nullptr;
    uint32_t numMaterials = 0;
    std::vector<int4> matTex;        // texture bindings per material (kept by set_materials)
    std::vector<void*> allocs;
    cx_scene_info info{};
};

struct cx_entries {
    cx_ctx* ctx = nullptr;
    uint32_t capacity = 0;
    EntryBuild E{};
    EntryTree T{};
    void* sortTemp = nullptr;
    size_t sortTempBytes = 0;
    std::vector<void*> allocs;
};

struct cx_renderer {
    cx_ctx* ctx = nullptr;
    const cx_scene* scene = nullptr;
    uint32_t W = 0, H = 0, P = 0;
    cx_launch_params params{};
    uint64_t frameRngState = 0;
    uint64_t frameIdx = 0;
    uint32_t front = 0, back = 0;
    uint4* prefixRes[2] = {nullptr, nullptr};
    uint4* suffixRes[2] = {nullptr, nullptr};
    uint4* gbuf = nullptr;
    uint32_t* neighIdx = nullptr;
    float* neighDist = nullptr;
    uint32_t* neighCount = nullptr;
    uint32_t kCap = 0;
    unsigned long long* counters = nullptr;
    cx_entries* entries = nullptr;
    cx_camera cam{};
    float* out = nullptr;
    bool timing = false;
    hipEvent_t ev[CX_PASS_COUNT + 1] = {};
    bool evMade = false;
    bool evValid = false;
    bool passRan[CX_PASS_COUNT] = {};
    // Event pairs around each traversal launch of the frame (timing enabled, wavefront mode).
    std::vector<hipEvent_t> traceEv;
    uint32_t traceEvUsed = 0;
    bool traceEvOk = true;
    std::vector<void*> allocs;
    int exec = CX_EXEC_WAVEFRONT;
    // Wavefront execution: two queues sharing one allocation each (viewed with the capacity of
    // the machine in flight) and per-pixel scratch.
    struct {
        uint4* ctx[2] = {nullptr, nullptr};
        uint32_t* job[2] = {nullptr, nullptr};
        float4* ray[2] = {nullptr, nullptr};
        uint4* res[2] = {nullptr, nullptr};
        uint32_t* count = nullptr;   // [2][2] closest / shadow count of each queue
        float4* shiftRes = nullptr;
        size_t jobCap = 0;           // jobs a queue holds
        size_t ctxWords = 0;         // 16-byte context words a queue holds
        uint4* prefixScratch = nullptr;
        uint4* suffixScratch = nullptr;
        uint4* packedScratch = nullptr;
        uint4* suffixAos = nullptr;   // record-major copy of the front suffix reservoirs (final gather)
        uint64_t* rngScratch = nullptr;
        uint32_t* pixFlags = nullptr;
        uint32_t* validMask = nullptr;
        uint32_t* gatherKeys = nullptr;      // Morton-ordered querying pixels of the final gather
        uint32_t* gatherList = nullptr;
        uint32_t* gatherKeysAlt = nullptr;
        uint32_t* gatherListAlt = nullptr;
        uint32_t* gatherCount = nullptr;
        uint32_t* gatherSortTemp = nullptr;
        size_t gatherSortBytes = 0;
        float4* gatherPos = nullptr;
        uint32_t* nbr = nullptr;
        float4* radScratch = nullptr;
        float4* fgScratch = nullptr;
        float4* camDir = nullptr;
        std::vector<void*> queueAllocs;
        std::vector<void*> pixelAllocs;
    } wave;
    // Longest path and prefix lengths since the last reset: they bound the number of
    // trace/step rounds a machine can need (records of earlier frames take part in shifts).
    uint32_t maxLenSeen = 0;
    uint32_t maxPrefixSeen = 0;
    // Most blocks of the grid-stride queue kernels.  1080p Conditional ReSTIR frame: 2048 / 4096 /
    // 8192 / 16384 -> 9.53 / 8.59-8.62 / 8.45 / 8.47 ms; a machine of fewer jobs launches fewer.
    uint32_t gridLimit = 8192;
    // Set by cx_batch_frame while it enqueues this renderer's frame beside other views: the
    // traversal then keeps whole 32-ray chunks (wave_trace minChunk); results are unchanged.
    bool concurrent = false;
    uint32_t waveRounds = CX_WAVE_ROUNDS_DEFAULT;   // rounds before the tail kernel, every machine (0: all; default per machine)
    uint64_t allocEpoch = 0;     // bumped whenever a buffer a frame uses is reallocated
    // Graph mode: a frame is one hipGraphLaunch of a graph captured from the eager frame for
    // each front-buffer parity; the camera, random and frame index come from `dyn`.
    bool graphMode = false;
    bool capturing = false;
    struct {
        hipGraphExec_t exec = nullptr;
        std::vector<uint8_t> key;
    } graphs[2];
    FrameDyn* dyn = nullptr;
    bool deviceRandom = false;   // frames draw their random number from devRng on the device
    uint64_t* devRng = nullptr;
    hipStream_t gstream = nullptr;   // capture / replay stream for frames on the null stream
    uint32_t captures = 0;
    // Final-phase overlap (wavefront Conditional ReSTIR): the entry store and the entry-tree
    // build run on `side` beside the gather's canonical prefix and suffix walks, joined by an
    // event before the neighbour query.
    bool overlap = true;
    hipStream_t side = nullptr;       // non-blocking, default priority
    hipEvent_t evFork = nullptr, evQuery = nullptr;
    hipStream_t fgSide = nullptr;     // set while frame_body launches the final gather
    // The caller's stream is recording a graph (hipStreamIsCapturing): the frame forks no side
    // stream and nothing may allocate, create streams or events, or synchronise.
    bool callerCapture = false;
};

namespace {

int set_err(cx_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int hip_err(cx_ctx* ctx, hipError_t e, const char* what) {
    return set_err(ctx, e == hipErrorOutOfMemory ? CX_ERR_OUT_OF_MEMORY : CX_ERR_HIP,
                   std::string(what) + ": " + hipGetErrorString(e));
}

#define CX_HIP(ctx, call)                                  \
    do {                                                   \
        hipError_t _e = (call);                            \
        if (_e != hipSuccess) return hip_err(ctx, _e, #call); \
    } while (0)

template <typename T>
hipError_t dalloc(std::vector<void*>& allocs, T** p, size_t count) {
    void* q = nullptr;
    const size_t bytes = std::max<size_t>(count * sizeof(T), 16);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return e;
    allocs.push_back(q);
    *p = (T*)q;
    return hipSuccess;
}

void free_all(std::vector<void*>& allocs) {
    for (void* p : allocs) (void)hipFree(p);
    allocs.clear();
}

// PCG32 step and seeding on the host (include/util/random.h:52-68), for the per-frame random.
uint32_t pcg_next(uint64_t& s) {
    const uint64_t old = s;
    s = old * PCG_MULT + PCG_INC;
    const uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
    const uint32_t rot = (uint32_t)(old >> 59);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

uint64_t pcg_seed(uint64_t initstate) {
    uint64_t s = 0;
    pcg_next(s);
    s += initstate;
    pcg_next(s);
    return s;
}

int entries_alloc(cx_ctx* ctx, uint32_t capacity, cx_entries** out) {
    cx_entries* en = new (std::nothrow) cx_entries();
    if (!en) return set_err(ctx, CX_ERR_OUT_OF_MEMORY, "host allocation failed");
    en->ctx = ctx;
    en->capacity = capacity;
    const uint32_t n = std::max(1u, capacity);
    hipError_t e = hipSuccess;
    auto A = [&](hipError_t r) {
        if (e == hipSuccess) e = r;
    };
    A(dalloc(en->allocs, &en->E.pos, n));
    A(dalloc(en->allocs, &en->E.valid, n));
    A(dalloc(en->allocs, &en->E.keys, n));
    A(dalloc(en->allocs, &en->E.vals, n));
    A(dalloc(en->allocs, &en->E.keysAlt, n));
    A(dalloc(en->allocs, &en->E.valsAlt, n));
    A(dalloc(en->allocs, &en->E.numValid, 1));
    A(dalloc(en->allocs, &en->T.entrySorted, n));
    A(dalloc(en->allocs, &en->T.entryPixel, n));
    A(dalloc(en->allocs, &en->T.nodes, 4 * (size_t)n));
    en->T.segLevelsCap = entry_seg_levels(n);
    const size_t segNodes = (size_t(2) << en->T.segLevelsCap) - 1;
    A(dalloc(en->allocs, &en->T.segLo, segNodes));
    A(dalloc(en->allocs, &en->T.segHi, segNodes));
    A(dalloc(en->allocs, &en->T.segLevels, 1));
    en->T.numValid = en->E.numValid;
    en->T.sortedKeys = en->E.keys;   // the sort leaves its output in keys / vals
    en->sortTempBytes = sort_temp_bytes(n);
    if (e == hipSuccess) {
        e = hipMalloc(&en->sortTemp, std::max<size_t>(en->sortTempBytes, 16));
        if (e == hipSuccess) en->allocs.push_back(en->sortTemp);
    }
    if (e == hipSuccess) e = hipMemset(en->E.numValid, 0, sizeof(uint32_t));
    if (e != hipSuccess) {
        free_all(en->allocs);
        delete en;
        return hip_err(ctx, e, "entries allocation");
    }
    *out = en;
    return CX_OK;
}

FrameArgs make_args(cx_renderer* r) {
    FrameArgs A;
    std::memset(&A, 0, sizeof(A));
    A.S = r->scene->d;
    A.lp = r->params;
    A.camPos = V3{r->cam.pos[0], r->cam.pos[1], r->cam.pos[2]};
    A.camU = V3{r->cam.U[0], r->cam.U[1], r->cam.U[2]};
    A.camV = V3{r->cam.V[0], r->cam.V[1], r->cam.V[2]};
    A.camW = V3{r->cam.W[0], r->cam.W[1], r->cam.W[2]};
    A.W = r->W;
    A.H = r->H;
    A.P = r->P;
    A.front = r->front;
    A.back = r->back;
    const cx_launch_params& lp = r->params;
    A.copyPrefixToBack = ((lp.rendererType == CX_RESTIR_PT && lp.prefixEnableSpatial) ||
                          (lp.rendererType == CX_CONDITIONAL_RESTIR && lp.suffixEnableSpatial))
                             ? 1u
                             : 0u;
    for (int b = 0; b < 2; b++) {
        A.prefixRes[b] = r->prefixRes[b];
        A.suffixRes[b] = r->suffixRes[b];
    }
    A.gbuf = r->gbuf;
    A.out = r->out;
    A.tree = r->entries->T;
    A.neighIdx = r->neighIdx;
    A.neighDist = r->neighDist;
    A.neighCount = r->neighCount;
    A.k = lp.gatherM > 1 ? (uint32_t)(lp.gatherM - 1) : 0u;
    A.counters = r->counters;
    A.dyn = (r->capturing || r->deviceRandom) ? r->dyn : nullptr;
    return A;
}

int check_params(cx_renderer* r, const cx_launch_params& p) {
    if (p.rendererType < 0 || p.rendererType > 2) return set_err(r->ctx, CX_ERR_INVALID_ARG, "bad rendererType");
    if (p.prefixSpatialCount < 0 || p.prefixSpatialCount > 16 || p.suffixSpatialCount < 0 ||
        p.suffixSpatialCount > 16)
        return set_err(r->ctx, CX_ERR_INVALID_ARG, "spatial neighbour count must be in [0, 16] (settings.h:24)");
    if (p.gatherM < 1 || p.gatherM > 32)
        return set_err(r->ctx, CX_ERR_INVALID_ARG, "gatherM must be in [1, 32] (Renderer.cpp:462)");
    if (p.gatherN < 1) return set_err(r->ctx, CX_ERR_INVALID_ARG, "gatherN must be >= 1");
    if (p.maxPathLen < 0 || p.prefixLen < 0 || p.maxPathLen > 255)
        return set_err(r->ctx, CX_ERR_INVALID_ARG, "bad path length");
    if (p.rendererType == CX_CONDITIONAL_RESTIR && p.prefixLen > p.maxPathLen)
        return set_err(r->ctx, CX_ERR_INVALID_ARG, "prefixLen must not exceed maxPathLen");
    if (p.windowRadius > 1024) return set_err(r->ctx, CX_ERR_INVALID_ARG, "windowRadius too large");
    return CX_OK;
}

int ensure_neighbors(cx_renderer* r, uint32_t k) {
    if (k <= r->kCap) return CX_OK;
    if (r->callerCapture)
        return set_err(r->ctx, CX_ERR_INVALID_ARG,
                       "the neighbour buffers would grow inside the caller's stream capture: run one frame "
                       "with these parameters eagerly first");
    if (r->neighIdx) {
        (void)hipFree(r->neighIdx);
        (void)hipFree(r->neighDist);
        r->allocs.erase(std::remove(r->allocs.begin(), r->allocs.end(), (void*)r->neighIdx), r->allocs.end());
        r->allocs.erase(std::remove(r->allocs.begin(), r->allocs.end(), (void*)r->neighDist), r->allocs.end());
        r->neighIdx = nullptr;
        r->neighDist = nullptr;
    }
    CX_HIP(r->ctx, dalloc(r->allocs, &r->neighIdx, (size_t)r->P * k));
    CX_HIP(r->ctx, dalloc(r->allocs, &r->neighDist, (size_t)r->P * k));
    r->kCap = k;
    r->allocEpoch++;
    return CX_OK;
}

void rec(cx_renderer* r, int i, hipStream_t s) {
    if (r->timing) (void)hipEventRecord(r->ev[i], s);
}

// ---- wavefront execution ----

uint32_t prefix_max_len(const cx_launch_params& lp) {
    return lp.rendererType == CX_RESTIR_PT ? (uint32_t)lp.maxPathLen : (uint32_t)lp.prefixLen;
}

// Shift jobs one pass of these parameters can create (two per pixel and neighbour slot).
size_t shift_jobs(const cx_renderer* r, const cx_launch_params& lp) {
    const uint32_t k = lp.gatherM > 1 ? (uint32_t)(lp.gatherM - 1) : 0u;
    const uint32_t slots = std::max({1u, (uint32_t)lp.prefixSpatialCount, (uint32_t)lp.suffixSpatialCount, k});
    return (size_t)2 * slots * r->P;
}

// Two queues' stripe counters, then the shared overflow flag.
constexpr size_t COUNT_WORDS = 2 * QSTRIPES * QCOUNT_STRIDE + 16;

int ensure_wave(cx_renderer* r, const cx_launch_params& lp) {
    auto& w = r->wave;
    hipError_t e = hipSuccess;
    auto A = [&](hipError_t x) {
        if (e == hipSuccess) e = x;
    };
    if (w.pixelAllocs.empty()) {
        if (r->callerCapture)
            return set_err(r->ctx, CX_ERR_INVALID_ARG,
                           "the wavefront buffers would be allocated inside the caller's stream capture: run one "
                           "frame eagerly first");
        const size_t P = r->P;
        A(dalloc(w.pixelAllocs, &w.prefixScratch, (size_t)PREFIX_CHUNKS * P));
        A(dalloc(w.pixelAllocs, &w.suffixScratch, (size_t)SUFFIX_CHUNKS * P));
        A(dalloc(w.pixelAllocs, &w.packedScratch, 2 * P));
        A(dalloc(w.pixelAllocs, &w.suffixAos, (size_t)SUFFIX_AOS_STRIDE * P));
        A(dalloc(w.pixelAllocs, &w.rngScratch, P));
        A(dalloc(w.pixelAllocs, &w.pixFlags, P));
        A(dalloc(w.pixelAllocs, &w.validMask, P));
        A(dalloc(w.pixelAllocs, &w.gatherKeys, P));
        A(dalloc(w.pixelAllocs, &w.gatherList, P));
        A(dalloc(w.pixelAllocs, &w.gatherKeysAlt, P));
        A(dalloc(w.pixelAllocs, &w.gatherListAlt, P));
        A(dalloc(w.pixelAllocs, &w.gatherCount, 1));
        w.gatherSortBytes = sort_temp_bytes((uint32_t)P);
        A(dalloc(w.pixelAllocs, &w.gatherSortTemp, w.gatherSortBytes / sizeof(uint32_t)));
        A(dalloc(w.pixelAllocs, &w.gatherPos, P));
        A(dalloc(w.pixelAllocs, &w.nbr, 16 * P));
        A(dalloc(w.pixelAllocs, &w.radScratch, P));
        A(dalloc(w.pixelAllocs, &w.fgScratch, P));
        A(dalloc(w.pixelAllocs, &w.camDir, P));
        A(dalloc(w.pixelAllocs, &w.count, COUNT_WORDS));
        if (e == hipSuccess) e = hipMemset(w.count, 0, COUNT_WORDS * sizeof(uint32_t));
        if (e != hipSuccess) {
            free_all(w.pixelAllocs);
            return hip_err(r->ctx, e, "wavefront scratch allocation");
        }
    }
    // slots of the striped queues of the pixel machines and of the shift machines
    const size_t capP = (size_t)QSTRIPES * queue_stripe_cap(r->P, r->P, r->gridLimit);
    const size_t capShift = (size_t)QSTRIPES * queue_stripe_cap((uint32_t)shift_jobs(r, lp), r->P, r->gridLimit);
    const size_t jobs = std::max(capP, capShift);
    // the largest context of the per-pixel path machines
    constexpr size_t pathChunks = std::max({PG_CHUNKS, SG_CHUNKS, CP_CHUNKS, FG_CHUNKS});
    const size_t words = std::max<size_t>(pathChunks * capP, (size_t)SH_CHUNKS * capShift);
    if (jobs <= w.jobCap && words <= w.ctxWords) return CX_OK;
    if (r->callerCapture)
        return set_err(r->ctx, CX_ERR_INVALID_ARG,
                       "the wavefront queues would grow inside the caller's stream capture: run one frame with "
                       "these parameters eagerly first");
    if (!w.queueAllocs.empty()) {
        e = hipDeviceSynchronize();
        if (e != hipSuccess) return hip_err(r->ctx, e, "wavefront queue reallocation");
        free_all(w.queueAllocs);
    }
    w.jobCap = 0;
    w.ctxWords = 0;
    for (int q = 0; q < 2; q++) {
        A(dalloc(w.queueAllocs, &w.ctx[q], words));
        A(dalloc(w.queueAllocs, &w.job[q], jobs));
        A(dalloc(w.queueAllocs, &w.ray[q], 2 * jobs));
        A(dalloc(w.queueAllocs, &w.res[q], jobs));
    }
    A(dalloc(w.queueAllocs, &w.shiftRes, jobs));
    if (e != hipSuccess) {
        free_all(w.queueAllocs);
        return hip_err(r->ctx, e, "wavefront queue allocation");
    }
    w.jobCap = jobs;
    w.ctxWords = words;
    r->allocEpoch++;
    return CX_OK;
}

// A view of queue `side` for a machine of at most `jobs` jobs (ensure_wave allocated every
// machine's striped size).
Queue queue_view(const cx_renderer* r, int side, uint32_t jobs) {
    Queue q;
    q.ctx = r->wave.ctx[side];
    q.job = r->wave.job[side];
    q.ray = r->wave.ray[side];
    q.res = r->wave.res[side];
    q.count = r->wave.count + (size_t)side * QSTRIPES * QCOUNT_STRIDE;
    q.overflow = r->wave.count + (size_t)2 * QSTRIPES * QCOUNT_STRIDE;
    q.jobs = jobs;
    q.maxBlocks = r->gridLimit;
    q.stripeCap = queue_stripe_cap(jobs, r->P, r->gridLimit);
    q.cap = QSTRIPES * q.stripeCap;
    return q;
}

WaveArgs make_wave_args(const cx_renderer* r, const FrameArgs& A) {
    WaveArgs W;
    std::memset(&W, 0, sizeof(W));
    W.A = A;
    const auto& w = r->wave;
    for (int b = 0; b < 2; b++) {
        W.R.prefixRes[b] = A.prefixRes[b];
        W.R.suffixRes[b] = A.suffixRes[b];
    }
    W.R.prefixScratch = w.prefixScratch;
    W.R.suffixScratch = w.suffixScratch;
    W.R.packedScratch = w.packedScratch;
    W.R.front = A.front;
    W.R.back = A.back;
    W.R.P = A.P;
    W.prefixScratch = w.prefixScratch;
    W.suffixScratch = w.suffixScratch;
    W.packedScratch = w.packedScratch;
    W.suffixAos = w.suffixAos;
    W.rngScratch = w.rngScratch;
    W.shiftRes = w.shiftRes;
    W.pixFlags = w.pixFlags;
    W.validMask = w.validMask;
    W.gatherKeys = w.gatherKeys;
    W.gatherList = w.gatherList;
    W.gatherKeysAlt = w.gatherKeysAlt;
    W.gatherListAlt = w.gatherListAlt;
    W.gatherCount = w.gatherCount;
    W.gatherSortTemp = w.gatherSortTemp;
    W.gatherSortBytes = w.gatherSortBytes;
    W.gatherPos = w.gatherPos;
    {   // the query points are quantised over the scene's bounds, as the entries are
        EntryBuild q{};
        set_quantisation(q, r->scene->info.boundsMin, r->scene->info.boundsMax);
        W.mortonLo = q.lo;
        W.mortonScale = q.scale;
    }
    W.nbr = w.nbr;
    W.radScratch = w.radScratch;
    W.fgScratch = w.fgScratch;
    W.camDir = w.camDir;
    return W;
}

// Runs one machine: `setup` (one thread per pixel) appends the first rays to queue 0, then
// `rounds` rounds of traversal + step.  Every job finishes within `rounds` rounds (the bounds
// below are the longest chains of trace calls of each machine), and a round whose queue is
// empty costs two near-empty launches.
// Records one event of the traversal timing pairs (grows the pool outside any capture).
void trace_rec(cx_renderer* r, hipStream_t s) {
    if (!r->timing) return;
    if (r->traceEvUsed == r->traceEv.size()) {
        hipEvent_t ev = nullptr;
        if (hipEventCreateWithFlags(&ev, hipEventDisableSystemFence) != hipSuccess) {
            r->traceEvOk = false;
            return;
        }
        r->traceEv.push_back(ev);
    }
    (void)hipEventRecord(r->traceEv[r->traceEvUsed++], s);
}

// Wavefront rounds before the tail kernel, per machine: 3 for suffix walks, 2 for the others.
// Measured on the 1080p Conditional ReSTIR frame (tent of traced lights, 14.6 rays/px): every
// machine 1 (suffix walks) or 2 (the others) / 2 / 3 / 4 / 6 rounds, or all rounds as traversal
// + step (0) -> 29.28 / 28.77 / 28.91 / 29.15 / 29.82 / 29.88 ms per frame; then, with the
// range-per-wavefront traversal, suffix walks 3 / 4 / 5 rounds against 2 -> 25.43 / 25.69 / 25.98
// against 25.77 ms, shifts 1 or 3 and prefix walks 1 or 3 against 2 no better.  The tail traces
// in place at 3 waves per SIMD, so long suffix walks are cheaper in the queue rounds.
uint32_t default_wave_rounds(int machine) { return machine == WK_SG_STEP ? 3u : 2u; }

hipError_t run_machine(cx_renderer* r, WaveArgs W, int setup, int step, uint32_t cap, uint32_t rounds,
                       int pass, hipStream_t s) {
    W.qin = queue_view(r, 1, cap);
    W.qout = queue_view(r, 0, cap);
    if ((size_t)W.qout.cap > r->wave.jobCap) return hipErrorInvalidValue;   // ensure_wave sized every machine
    hipError_t e = wave_zero_counts(W.qout.count, s);
    if (e == hipSuccess) e = wave_launch(setup, W, r->P, s);
    int cur = 0;
    const uint32_t wr = r->waveRounds != CX_WAVE_ROUNDS_DEFAULT ? r->waveRounds : default_wave_rounds(step);
    const uint32_t waveRounds = (wr > 0 && wr < rounds) ? wr : rounds;
    for (uint32_t i = 0; i < waveRounds && e == hipSuccess; i++) {
        const Queue qin = queue_view(r, cur, cap), qout = queue_view(r, 1 - cur, cap);
        trace_rec(r, s);
        e = wave_trace(W.A.S, qin, qout, W.A.counters + 2 * pass, r->concurrent ? ~0u : 1u, s);
        trace_rec(r, s);
        W.qin = qin;
        W.qout = qout;
        if (e == hipSuccess) e = wave_launch(step, W, 0, s);
        cur = 1 - cur;
    }
    if (e == hipSuccess && waveRounds < rounds) {   // the remaining rounds in one launch
        W.qin = queue_view(r, cur, cap);
        W.qout = queue_view(r, 1 - cur, cap);
        e = wave_tail(step, W, rounds - waveRounds, pass, s);
    }
    return e;
}

// A pass of the frame loop in wavefront form; same results as launch_pass.
hipError_t wave_pass(cx_renderer* r, int pass, const FrameArgs& A, hipStream_t s) {
    const cx_launch_params& lp = A.lp;
    WaveArgs W = make_wave_args(r, A);
    const uint32_t P = r->P;
    const uint32_t shiftCap = (uint32_t)shift_jobs(r, lp);
    const uint32_t L = std::max(r->maxLenSeen, (uint32_t)lp.maxPathLen);
    const uint32_t PL = std::max(r->maxPrefixSeen, prefix_max_len(lp));
    // Longest trace chains, each round answering one ray per job:
    //  * prefix walk: up to PL closest hits + 1 shadow ray;
    //  * complete path: maxPathLen closest hits + 1 shadow ray;
    //  * suffix walk: a prefix of length >= 1 leaves at most L - 1 vertices, the first one
    //    rehydrated, so up to L - 2 closest hits + 1 shadow + 1 reconnection ray;
    //  * prefix shift: reconIdx <= PL, replays reconIdx - 2 vertices + 1 shadow ray;
    //  * suffix shift: reconIdx <= max(PL, L - 1) (path_gen.h:338 and :378), replays
    //    reconIdx - 1 vertices + 1 shadow ray;
    //  * final-gather prefix: prefixLen closest hits.
    const uint32_t pgRounds = PL + 1, cpRounds = (uint32_t)lp.maxPathLen + 1;
    const uint32_t sgRounds = std::max(L, 2u);
    const uint32_t shPrefixRounds = std::max(PL, 2u) - 1;
    const uint32_t shSuffixRounds = std::max({PL, L > 0 ? L - 1 : 0u, 1u});
    const uint32_t fgRounds = (uint32_t)lp.prefixLen;
    const bool restir = lp.rendererType != CX_PATH_TRACER;
    hipError_t e = hipSuccess;
    switch (pass) {
        case CX_PASS_PREFIX_GEN_TEMP:
            if (!restir) return run_machine(r, W, WK_PREFIX_INIT, WK_CP_STEP, P, cpRounds, pass, s);
            e = run_machine(r, W, WK_PREFIX_INIT, WK_PG_STEP, P, pgRounds, pass, s);
            if (e == hipSuccess)
                e = run_machine(r, W, WK_PREFIX_TEMPORAL_SETUP, WK_SH_STEP, shiftCap,
                                lp.prefixEnableTemporal ? shPrefixRounds : 0u, pass, s);
            if (e == hipSuccess) e = wave_launch(WK_PREFIX_RESAMPLE, W, P, s);
            return e;
        case CX_PASS_PREFIX_SPATIAL:
            W.S = (uint32_t)lp.prefixSpatialCount;
            e = run_machine(r, W, WK_PREFIX_SPATIAL_SETUP, WK_SH_STEP, shiftCap, shPrefixRounds, pass, s);
            if (e == hipSuccess) e = wave_launch(WK_PREFIX_SPATIAL_RESAMPLE, W, P, s);
            return e;
        case CX_PASS_SUFFIX_GEN_TEMP:
            e = run_machine(r, W, WK_SUFFIX_INIT, WK_SG_STEP, P, sgRounds, pass, s);
            if (e == hipSuccess)
                e = run_machine(r, W, WK_SUFFIX_TEMPORAL_SETUP, WK_SH_STEP, shiftCap,
                                lp.suffixEnableTemporal ? shSuffixRounds : 0u, pass, s);
            if (e == hipSuccess) e = wave_launch(WK_SUFFIX_RESAMPLE, W, P, s);
            return e;
        case CX_PASS_SUFFIX_SPATIAL:
            W.S = (uint32_t)lp.suffixSpatialCount;
            e = run_machine(r, W, WK_SUFFIX_SPATIAL_SETUP, WK_SH_STEP, shiftCap, shSuffixRounds, pass, s);
            if (e == hipSuccess) e = wave_launch(WK_SUFFIX_SPATIAL_RESAMPLE, W, P, s);
            return e;
        case CX_PASS_FINAL_GATHER: {
            const bool more = A.k > 0 && lp.gatherN > 1;
            hipStream_t side = r->fgSide;
            // the neighbours' records, read at random by the shifts and the combination
            e = wave_launch(WK_SUFFIX_AOS, W, P * SUFFIX_AOS_STRIDE, s);
            W.R.suffixFrontAos = W.suffixAos;
            if (e == hipSuccess) e = run_machine(r, W, WK_GATHER_INIT, WK_PG_STEP, P, pgRounds, pass, s);
            if (e == hipSuccess) e = run_machine(r, W, WK_GATHER_SUFFIX_INIT, WK_SG_STEP, P, sgRounds, pass, s);
            // the entry store and tree build, forked to `side`, join before the query
            if (e == hipSuccess && side) {
                e = hipEventRecord(r->evQuery, side);
                if (e == hipSuccess) e = hipStreamWaitEvent(s, r->evQuery, 0);
            }
            W.gIdx = 0;
            if (e == hipSuccess) e = run_machine(r, W, WK_GATHER_SETUP, WK_SH_STEP, shiftCap, shSuffixRounds, pass, s);
            if (e == hipSuccess) e = wave_launch(more ? WK_GATHER_COMBINE : WK_GATHER_COMBINE_LAST, W, P, s);
            for (int g = 1; more && g < lp.gatherN && e == hipSuccess; g++) {
                W.gIdx = (uint32_t)g;
                e = run_machine(r, W, WK_FG_INIT, WK_FG_STEP, P, fgRounds, pass, s);
                if (e == hipSuccess) e = run_machine(r, W, WK_GATHER_SETUP, WK_SH_STEP, shiftCap, shSuffixRounds, pass, s);
                if (e == hipSuccess)
                    e = wave_launch(g == lp.gatherN - 1 ? WK_GATHER_COMBINE_LAST : WK_GATHER_COMBINE, W, P, s);
            }
            return e;
        }
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

namespace {

// Whether the caller's stream is recording a graph.  Inside the library's own graph-mode capture
// (r->capturing) the frame is not the caller's; a stream whose capture was invalidated is
// refused.
int detect_caller_capture(cx_renderer* r, void* stream) {
    if (r->capturing) return CX_OK;
    r->callerCapture = false;
    if (!stream) return CX_OK;   // the legacy null stream cannot be captured
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    const hipError_t e = hipStreamIsCapturing((hipStream_t)stream, &cs);
    if (e != hipSuccess) return hip_err(r->ctx, e, "hipStreamIsCapturing");
    if (cs == hipStreamCaptureStatusInvalidated)
        return set_err(r->ctx, CX_ERR_INVALID_ARG, "the caller's stream capture has been invalidated");
    r->callerCapture = cs == hipStreamCaptureStatusActive;
    return CX_OK;
}

}  // namespace

namespace cx {
void ctx_set_error(cx_ctx* ctx, const char* msg) {
    if (ctx) ctx->err = msg;
}
int ctx_device(const cx_ctx* ctx) { return ctx ? ctx->device : -1; }
}  // namespace cx

extern "C" {

int cx_abi_version(void) { return CX_ABI_VERSION; }

int cx_ctx_create(int device, cx_ctx** out) {
    if (!out) return CX_ERR_INVALID_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return CX_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return CX_ERR_INVALID_ARG;
    if (hipSetDevice(device) != hipSuccess) return CX_ERR_HIP;
    // the static scratch of cx_entries_search_static: its address on this device, taken once
    if (noalloc_setup(device) != hipSuccess) return CX_ERR_HIP;
    cx_ctx* c = new (std::nothrow) cx_ctx();
    if (!c) return CX_ERR_OUT_OF_MEMORY;
    c->device = device;
    *out = c;
    return CX_OK;
}

int cx_ctx_destroy(cx_ctx* ctx) {
    if (ctx)
        for (hipStream_t s : ctx->pool) (void)hipStreamDestroy(s);
    delete ctx;
    return CX_OK;
}

const char* cx_last_error(const cx_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cx_scene_build(cx_ctx* ctx, const cx_scene_desc* d, cx_scene** out) {
    if (!ctx || !d || !out) return set_err(ctx, CX_ERR_INVALID_ARG, "null argument");
    if (d->numTriangles > 0 && (!d->positions || !d->normals || !d->indices || !d->triMaterial))
        return set_err(ctx, CX_ERR_INVALID_ARG, "missing geometry arrays");
    if (d->numEmitters == 0 || !d->emitterFaces || !d->emitterFaceOffset || !d->emitterColor)
        return set_err(ctx, CX_ERR_INVALID_ARG, "at least one emitter is required (sample_emitter.h:5-6)");
    if (d->numMaterials == 0 || !d->materials) return set_err(ctx, CX_ERR_INVALID_ARG, "no materials");
    for (uint32_t t = 0; t < d->numTriangles; t++) {
        if (d->triMaterial[t] >= d->numMaterials) return set_err(ctx, CX_ERR_INVALID_ARG, "material id out of range");
        for (int k = 0; k < 3; k++)
            if (d->indices[3 * (size_t)t + k] >= d->numVertices)
                return set_err(ctx, CX_ERR_INVALID_ARG, "vertex index out of range");
    }
    const uint32_t numFaces = d->emitterFaceOffset[d->numEmitters];
    for (uint32_t e = 0; e < d->numEmitters; e++)
        if (d->emitterFaceOffset[e + 1] <= d->emitterFaceOffset[e])
            return set_err(ctx, CX_ERR_INVALID_ARG, "every emitter needs at least one face");
    // Textures: the sampler's wrap arithmetic needs 1 <= size <= 65536 per axis, and texel
    // offsets are 32-bit.
    if (d->numTextures && !d->textures) return set_err(ctx, CX_ERR_INVALID_ARG, "missing texture array");
    std::vector<uint4> texInfo(d->numTextures);
    uint64_t numTexels = 0;
    for (uint32_t t = 0; t < d->numTextures; t++) {
        const cx_texture& tx = d->textures[t];
        if (!tx.rgba || tx.width == 0 || tx.height == 0 || tx.width > 65536u || tx.height > 65536u)
            return set_err(ctx, CX_ERR_INVALID_ARG, "texture size must be 1..65536 per axis with data");
        texInfo[t] = make_uint4((uint32_t)numTexels, tx.width, tx.height, 0u);
        numTexels += (uint64_t)tx.width * tx.height;
        if (numTexels >= (1ull << 32)) return set_err(ctx, CX_ERR_INVALID_ARG, "more than 2^32 texels");
    }
    std::vector<int4> matTex(d->numMaterials, make_int4(-1, -1, -1, 0));
    if (d->materialTex)
        for (uint32_t m = 0; m < d->numMaterials; m++) {
            const cx_material_tex& mt = d->materialTex[m];
            const int32_t ids[3] = {mt.diffTex, mt.specTex, mt.roughTex};
            for (int k = 0; k < 3; k++)
                if (ids[k] < -1 || (ids[k] >= 0 && (uint32_t)ids[k] >= d->numTextures))
                    return set_err(ctx, CX_ERR_INVALID_ARG, "material texture index out of range");
            if (mt.specTexUsage > CX_SPEC_TEX_OCC_ROUGH_METAL)
                return set_err(ctx, CX_ERR_INVALID_ARG, "unknown specTexUsage");
            matTex[m] = make_int4(mt.diffTex, mt.specTex, mt.roughTex, (int32_t)mt.specTexUsage);
        }
    const uint32_t wantWidth = d->bvhWidth ? d->bvhWidth : 4u;
    if (wantWidth != 2 && wantWidth != 4 && wantWidth != 8)
        return set_err(ctx, CX_ERR_INVALID_ARG, "bvhWidth must be 0, 2, 4 or 8");
    if (d->bvhBuilder != CX_BVH_BUILD_HOST && d->bvhBuilder != CX_BVH_BUILD_DEVICE)
        return set_err(ctx, CX_ERR_INVALID_ARG, "unknown bvhBuilder");
    const bool onDevice = d->bvhBuilder == CX_BVH_BUILD_DEVICE;
    // Both builders compare box areas and centroids: they need finite coordinates.
    for (size_t i = 0; i < 3 * (size_t)d->numVertices; i++)
        if (!std::isfinite(d->positions[i])) return set_err(ctx, CX_ERR_INVALID_ARG, "non-finite vertex position");
    const uint32_t maxLeaf = d->bvhMaxLeafSize ? d->bvhMaxLeafSize : DEFAULT_MAX_LEAF;
    const auto tBuild = std::chrono::steady_clock::now();
    BvhHost bvh;
    std::string err;
    std::vector<float4> wide;
    int32_t root = 0, wroot = 0;
    uint32_t stackBound = 0, width = 2, maxDepth = 0, numNodes = 0, numWide = 0;
    if (!onDevice) {
        if (!build_bvh(d->positions, d->indices, d->numTriangles, maxLeaf, bvh, err, d->triMask))
            return set_err(ctx, CX_ERR_BUILD, err);
        root = wroot = bvh.root;
        stackBound = maxDepth = bvh.maxDepth;
        if (wantWidth > 2 && collapse_bvh(bvh, (int)wantWidth, (uint32_t)STACK_DEPTH, wide, wroot, stackBound))
            width = wantWidth;
        if (width == 2) stackBound = bvh.maxDepth;
        numNodes = (uint32_t)(bvh.nodes.size() / 4);
        numWide = width > 2 ? (uint32_t)(wide.size() / (2 * width)) : 0u;
    }
    cx_scene* s = new (std::nothrow) cx_scene();
    if (!s) return set_err(ctx, CX_ERR_OUT_OF_MEMORY, "host allocation failed");
    s->ctx = ctx;
    hipError_t e = hipSuccess;
    auto up = [&](const void* src, size_t count, size_t elem) -> void* {
        if (e != hipSuccess) return nullptr;
        void* q = nullptr;
        e = hipMalloc(&q, std::max<size_t>(count * elem, 16));
        if (e != hipSuccess) return nullptr;
        s->allocs.push_back(q);
        if (count) e = hipMemcpy(q, src, count * elem, hipMemcpyHostToDevice);
        return q;
    };
    std::vector<DMaterial> mats(d->numMaterials);
    for (uint32_t m = 0; m < d->numMaterials; m++) {
        const cx_material& cm = d->materials[m];
        mats[m].diffRough = make_float4(cm.diffColor[0], cm.diffColor[1], cm.diffColor[2], cm.roughness);
        mats[m].spec = make_float4(cm.specF0[0], cm.specF0[1], cm.specF0[2], 0.0f);
        mats[m].emissive = make_float4(cm.emissiveColor[0], cm.emissiveColor[1], cm.emissiveColor[2], 0.0f);
        mats[m].tex = matTex[m];
    }
    std::vector<uint32_t> texels((size_t)numTexels);
    for (uint32_t t = 0; t < d->numTextures; t++)
        std::memcpy(texels.data() + texInfo[t].x, d->textures[t].rgba, (size_t)texInfo[t].y * texInfo[t].z * 4);
    if (!onDevice) {
        const float4* nodesSrc = bvh.nodes.empty() ? nullptr : bvh.nodes.data();
        s->d.nodes = (const float4*)up(nodesSrc, bvh.nodes.size(), sizeof(float4));
        s->d.wnodes = width > 2 ? (const float4*)up(wide.data(), wide.size(), sizeof(float4)) : nullptr;
        s->d.tris = (const float4*)up(bvh.tris.data(), bvh.tris.size(), sizeof(float4));
    }
    s->d.pos = (const float*)up(d->positions, 3 * (size_t)d->numVertices, sizeof(float));
    s->d.nrm = (const float*)up(d->normals, 3 * (size_t)d->numVertices, sizeof(float));
    s->d.idx = (const uint32_t*)up(d->indices, 3 * (size_t)d->numTriangles, sizeof(uint32_t));
    s->d.triMat = (const uint32_t*)up(d->triMaterial, d->numTriangles, sizeof(uint32_t));
    s->d.mats = (const DMaterial*)up(mats.data(), mats.size(), sizeof(DMaterial));
    s->evalMats = (DMaterial*)up(mats.data(), mats.size(), sizeof(DMaterial));
    s->numMaterials = d->numMaterials;
    s->d.emitFaces = (const float*)up(d->emitterFaces, 9 * (size_t)numFaces, sizeof(float));
    s->d.emitOff = (const uint32_t*)up(d->emitterFaceOffset, d->numEmitters + 1, sizeof(uint32_t));
    s->d.emitColor = (const float*)up(d->emitterColor, 3 * (size_t)d->numEmitters, sizeof(float));
    s->evalEmit = (float*)up(d->emitterColor, 3 * (size_t)d->numEmitters, sizeof(float));
    s->d.uv = d->uvs ? (const float*)up(d->uvs, 2 * (size_t)d->numVertices, sizeof(float)) : nullptr;
    s->d.texels = (const uint32_t*)up(texels.data(), texels.size(), sizeof(uint32_t));
    s->d.texInfo = (const uint4*)up(texInfo.data(), texInfo.size(), sizeof(uint4));
    s->d.numTextures = d->numTextures;
    s->matTex = matTex;
    if (e != hipSuccess) {
        free_all(s->allocs);
        delete s;
        return hip_err(ctx, e, "scene upload");
    }
    uint32_t builderUsed = d->bvhBuilder;
    if (onDevice) {
        DeviceBvh db;
        hipError_t he = hipSuccess;
        std::vector<void*> bvhAllocs;
        const uint8_t* dmask =
            d->triMask ? (const uint8_t*)up(d->triMask, d->numTriangles, sizeof(uint8_t)) : nullptr;
        if (e != hipSuccess) {
            free_all(s->allocs);
            delete s;
            return hip_err(ctx, e, "scene upload");
        }
        if (!build_bvh_device(s->d.pos, s->d.idx, dmask, d->numTriangles, maxLeaf, (int)wantWidth, db, bvhAllocs, err,
                              he)) {
            free_all(bvhAllocs);
            if (!db.tooDeep) {
                free_all(s->allocs);
                delete s;
                if (he != hipSuccess && he != hipErrorUnknown) return hip_err(ctx, he, err.c_str());
                return set_err(ctx, CX_ERR_BUILD, err);
            }
            // a tree deeper than the traversal stack (heavily duplicated geometry): the host
            // build bounds its depth, so it takes over; cx_scene_info.bvhBuilder reports it
            builderUsed = CX_BVH_BUILD_HOST;
            if (!build_bvh(d->positions, d->indices, d->numTriangles, maxLeaf, bvh, err, d->triMask)) {
                free_all(s->allocs);
                delete s;
                return set_err(ctx, CX_ERR_BUILD, err);
            }
            db = DeviceBvh{};
            db.root = db.wroot = bvh.root;
            db.stackBound = db.maxDepth = bvh.maxDepth;
            if (wantWidth > 2 && collapse_bvh(bvh, (int)wantWidth, (uint32_t)STACK_DEPTH, wide, db.wroot, db.stackBound))
                db.width = wantWidth;
            if (db.width == 2) db.stackBound = bvh.maxDepth;
            db.numNodes = (uint32_t)(bvh.nodes.size() / 4);
            db.numWide = db.width > 2 ? (uint32_t)(wide.size() / (2 * db.width)) : 0u;
            db.nodes = (float4*)up(bvh.nodes.empty() ? nullptr : bvh.nodes.data(), bvh.nodes.size(), sizeof(float4));
            db.wnodes = db.width > 2 ? (float4*)up(wide.data(), wide.size(), sizeof(float4)) : nullptr;
            db.tris = (float4*)up(bvh.tris.data(), bvh.tris.size(), sizeof(float4));
        } else {
            s->allocs.insert(s->allocs.end(), bvhAllocs.begin(), bvhAllocs.end());
        }
        s->d.nodes = db.nodes ? db.nodes : (const float4*)up(nullptr, 0, sizeof(float4));
        s->d.tris = db.tris ? db.tris : (const float4*)up(nullptr, 0, sizeof(float4));
        s->d.wnodes = db.wnodes;
        if (db.width > 2 && !db.wnodes) s->d.wnodes = (const float4*)up(nullptr, 0, sizeof(float4));
        if (e != hipSuccess) {
            free_all(s->allocs);
            delete s;
            return hip_err(ctx, e, "scene upload");
        }
        root = db.root;
        wroot = db.wroot;
        width = db.width;
        maxDepth = db.maxDepth;
        stackBound = db.stackBound;
        numNodes = db.numNodes;
        numWide = db.numWide;
    }
    const float buildMs =
        std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - tBuild).count();
    s->d.numTris = d->numTriangles;
    s->d.numEmitters = d->numEmitters;
    s->d.root = root;
    s->d.wroot = wroot;
    s->d.width = width;
    // Scene::GetSceneSize (src/Scene.cpp:48-71): bbox of all vertices and the origin.
    float mn[3] = {0.0f, 0.0f, 0.0f}, mx[3] = {0.0f, 0.0f, 0.0f};
    for (uint32_t v = 0; v < d->numVertices; v++)
        for (int a = 0; a < 3; a++) {
            const float c = d->positions[3 * (size_t)v + a];
            mn[a] = (c < mn[a]) ? c : mn[a];
            mx[a] = (mx[a] < c) ? c : mx[a];
        }
    const float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    s->info.sceneSize = std::sqrt(dx * dx + dy * dy + dz * dz);
    for (int a = 0; a < 3; a++) {
        s->info.boundsMin[a] = mn[a];
        s->info.boundsMax[a] = mx[a];
    }
    s->info.numTriangles = d->numTriangles;
    s->info.numVertices = d->numVertices;
    s->info.numNodes = numNodes;
    s->info.maxDepth = maxDepth;
    s->info.numEmitters = d->numEmitters;
    s->info.numEmitterFaces = numFaces;
    s->info.bvhWidth = width;
    s->info.numWideNodes = numWide;
    s->info.stackBound = stackBound;
    s->info.numTextures = d->numTextures;
    s->info.bvhBuilder = builderUsed;
    s->info.buildMs = buildMs;
    s->info.deviceBytes = (4 * (size_t)numNodes + 2 * (size_t)width * numWide + 3 * (size_t)d->numTriangles) *
                              sizeof(float4) +
                          (size_t)d->numVertices * (d->uvs ? 32 : 24) + (size_t)d->numTriangles * 16 +
                          mats.size() * sizeof(DMaterial) + texels.size() * 4 + texInfo.size() * sizeof(uint4);
    *out = s;
    return CX_OK;
}

int cx_scene_destroy(cx_scene* s) {
    if (!s) return CX_OK;
    free_all(s->allocs);
    delete s;
    return CX_OK;
}

int cx_scene_get_info(const cx_scene* s, cx_scene_info* info) {
    if (!s || !info) return CX_ERR_INVALID_ARG;
    *info = s->info;
    return CX_OK;
}

int cx_tonemap(cx_ctx* ctx, const float* hdr, uint8_t* ldr, uint32_t n, void* stream) {
    if (!ctx || (n && (!hdr || !ldr))) return set_err(ctx, CX_ERR_INVALID_ARG, "null argument");
    CX_HIP(ctx, launch_tonemap(hdr, ldr, n, (hipStream_t)stream));
    return CX_OK;
}

int cx_scene_set_materials(cx_scene* s, const cx_material* mats, uint32_t n, int tables) {
    if (!s || !mats) return CX_ERR_INVALID_ARG;
    if (n != s->numMaterials) return set_err(s->ctx, CX_ERR_INVALID_ARG, "material count differs from the scene's");
    if (tables & ~(CX_TABLE_SAMPLING | CX_TABLE_EVAL)) return set_err(s->ctx, CX_ERR_INVALID_ARG, "unknown table");
    std::vector<DMaterial> dm(n);
    for (uint32_t m = 0; m < n; m++) {
        const cx_material& cm = mats[m];
        dm[m].diffRough = make_float4(cm.diffColor[0], cm.diffColor[1], cm.diffColor[2], cm.roughness);
        dm[m].spec = make_float4(cm.specF0[0], cm.specF0[1], cm.specF0[2], 0.0f);
        dm[m].emissive = make_float4(cm.emissiveColor[0], cm.emissiveColor[1], cm.emissiveColor[2], 0.0f);
        dm[m].tex = s->matTex[m];
    }
    if (tables & CX_TABLE_SAMPLING)
        CX_HIP(s->ctx, hipMemcpy((void*)s->d.mats, dm.data(), n * sizeof(DMaterial), hipMemcpyHostToDevice));
    if (tables & CX_TABLE_EVAL)
        CX_HIP(s->ctx, hipMemcpy(s->evalMats, dm.data(), n * sizeof(DMaterial), hipMemcpyHostToDevice));
    return CX_OK;
}

int cx_scene_set_emitter_colors(cx_scene* s, const float* colors, uint32_t n, int tables) {
    if (!s || !colors) return CX_ERR_INVALID_ARG;
    if (n != s->d.numEmitters) return set_err(s->ctx, CX_ERR_INVALID_ARG, "emitter count differs from the scene's");
    if (tables & ~(CX_TABLE_SAMPLING | CX_TABLE_EVAL)) return set_err(s->ctx, CX_ERR_INVALID_ARG, "unknown table");
    if (tables & CX_TABLE_SAMPLING)
        CX_HIP(s->ctx, hipMemcpy((void*)s->d.emitColor, colors, 3 * (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    if (tables & CX_TABLE_EVAL)
        CX_HIP(s->ctx, hipMemcpy(s->evalEmit, colors, 3 * (size_t)n * sizeof(float), hipMemcpyHostToDevice));
    return CX_OK;
}

int cx_bvh_host_trace(const cx_scene_desc* d, const float* rays, uint32_t n, int anyHit, uint32_t* prim, float* t,
                      float* u, float* v, uint32_t* widthUsed) {
    if (!d || (n && (!rays || !prim || (!anyHit && (!t || !u || !v))))) return CX_ERR_INVALID_ARG;
    const uint32_t wantWidth = d->bvhWidth ? d->bvhWidth : 4u;
    if (wantWidth != 2 && wantWidth != 4 && wantWidth != 8) return CX_ERR_INVALID_ARG;
    for (size_t i = 0; i < 3 * (size_t)d->numVertices; i++)
        if (!std::isfinite(d->positions[i])) return CX_ERR_INVALID_ARG;
    BvhHost bvh;
    std::string err;
    if (!build_bvh(d->positions, d->indices, d->numTriangles, d->bvhMaxLeafSize ? d->bvhMaxLeafSize : DEFAULT_MAX_LEAF,
                   bvh, err, d->triMask))
        return CX_ERR_BUILD;
    std::vector<float4> wide;
    int32_t wroot = bvh.root;
    uint32_t stackBound = 0, width = 2;
    if (wantWidth > 2 && collapse_bvh(bvh, (int)wantWidth, (uint32_t)STACK_DEPTH, wide, wroot, stackBound))
        width = wantWidth;
    if (widthUsed) *widthUsed = width;
    DScene S{};
    S.nodes = bvh.nodes.data();
    S.wnodes = width > 2 ? wide.data() : nullptr;
    S.tris = bvh.tris.data();
    S.numTris = d->numTriangles;
    S.root = bvh.root;
    S.wroot = wroot;
    S.width = width;
    struct HostStack {
        int e[STACK_DEPTH];
        void put(int i, int x) { e[i] = x; }
        int get(int i) const { return e[i]; }
    };
    // The stepped traversal of the queue-trace kernel (trav_init / trav_step), one ray at a time.
    auto run = [&](auto widthTag, auto anyTag, TravRay& tr, HostStack& st) {
        constexpr int Wd = decltype(widthTag)::value;
        constexpr bool Any = decltype(anyTag)::value;
        while (!trav_step<Any, Wd>(S, tr, st)) {
        }
    };
    for (uint32_t i = 0; i < n; i++) {
        const float* r = rays + 8 * (size_t)i;
        HostStack st;
        TravRay tr;
        const V3 o{r[0], r[1], r[2]}, dir{r[3], r[4], r[5]};
        auto go = [&](auto anyTag) {
            if (width == 4) {
                trav_init<4>(S, tr, o, dir, r[6], r[7]);
                run(std::integral_constant<int, 4>{}, anyTag, tr, st);
            } else if (width == 8) {
                trav_init<8>(S, tr, o, dir, r[6], r[7]);
                run(std::integral_constant<int, 8>{}, anyTag, tr, st);
            } else {
                trav_init<2>(S, tr, o, dir, r[6], r[7]);
                run(std::integral_constant<int, 2>{}, anyTag, tr, st);
            }
        };
        if (anyHit) {
            go(std::true_type{});
            prim[i] = tr.best.prim != INVALID_PRIM ? 1u : 0u;
        } else {
            go(std::false_type{});
            const Hit& h = tr.best;
            const bool hit = h.prim != INVALID_PRIM;
            prim[i] = h.prim;
            t[i] = hit ? h.t : 0.0f;
            u[i] = hit ? h.u : 0.0f;
            v[i] = hit ? h.v : 0.0f;
        }
    }
    return CX_OK;
}

int cx_trace_closest(cx_ctx* ctx, const cx_scene* s, const float* rays, uint32_t n, uint32_t* prim, float* t,
                     float* u, float* v, void* stream) {
    if (!ctx || !s || (n && (!rays || !prim || !t || !u || !v))) return set_err(ctx, CX_ERR_INVALID_ARG, "null argument");
    CX_HIP(ctx, launch_trace(s->d, rays, n, 0, prim, t, u, v, nullptr, (hipStream_t)stream));
    return CX_OK;
}

int cx_trace_occluded(cx_ctx* ctx, const cx_scene* s, const float* rays, uint32_t n, uint32_t* occ, void* stream) {
    if (!ctx || !s || (n && (!rays || !occ))) return set_err(ctx, CX_ERR_INVALID_ARG, "null argument");
    CX_HIP(ctx, launch_trace(s->d, rays, n, 1, nullptr, nullptr, nullptr, nullptr, occ, (hipStream_t)stream));
    return CX_OK;
}

int cx_entries_create(cx_ctx* ctx, uint32_t capacity, cx_entries** out) {
    if (!ctx || !out) return set_err(ctx, CX_ERR_INVALID_ARG, "null argument");
    return entries_alloc(ctx, capacity, out);
}

int cx_entries_destroy(cx_entries* en) {
    if (!en) return CX_OK;
    free_all(en->allocs);
    delete en;
    return CX_OK;
}

int cx_entries_build(cx_ctx* ctx, cx_entries* en, const float* pos, const float* r, const uint8_t* valid, uint32_t n,
                     const float* sceneMin, const float* sceneMax, void* stream) {
    if (!ctx || !en || !sceneMin || !sceneMax || (n && (!pos || !r || !valid)))
        return set_err(ctx, CX_ERR_INVALID_ARG, "null argument");
    if (n > en->capacity) return set_err(ctx, CX_ERR_INVALID_ARG, "more entries than the structure's capacity");
    set_quantisation(en->E, sceneMin, sceneMax);
    hipStream_t s = (hipStream_t)stream;
    CX_HIP(ctx, launch_entries_from_arrays(pos, r, valid, n, en->E, s));
    CX_HIP(ctx, launch_tree_build(en->E, en->T, n, en->sortTemp, en->sortTempBytes, s));
    return CX_OK;
}

int cx_entries_query(cx_ctx* ctx, const cx_entries* en, const float* q, uint32_t nq, uint32_t k, uint32_t* idx,
                     float* dist, uint32_t* count, void* stream) {
    if (!ctx || !en || (nq && (!q || !idx || !dist || !count))) return set_err(ctx, CX_ERR_INVALID_ARG, "null argument");
    if (k == 0 || k > 31) return set_err(ctx, CX_ERR_INVALID_ARG, "k must be in [1, 31]");
    CX_HIP(ctx, launch_entries_query(en->T, q, nq, k, idx, dist, count, (hipStream_t)stream));
    return CX_OK;
}

int cx_default_params(const cx_scene* s, uint32_t w, uint32_t h, cx_launch_params* p) {
    if (!p) return CX_ERR_INVALID_ARG;
    std::memset(p, 0, sizeof(*p));
    p->width = w;
    p->height = h;
    p->windowRadius = 10;
    p->neeProb = 0.3f;
    p->maxPathLen = 8;
    p->rendererType = CX_PATH_TRACER;
    p->reconMinDistance = 0.01f * (s ? s->info.sceneSize : 0.0f);
    p->reconMinRoughness = 0.2f;
    p->prefixLen = 2;
    p->prefixEnableTemporal = 1;
    p->prefixEnableSpatial = 0;
    p->prefixSpatialCount = 1;
    p->prefixSpatialRounds = 1;
    p->suffixEnableTemporal = 1;
    p->suffixEnableSpatial = 1;
    p->suffixSpatialCount = 1;
    p->suffixSpatialRounds = 1;
    p->gatherN = 1;
    p->gatherM = 1;
    p->gatherRadius = 0.01f;
    p->gatherRadiusType = CX_RADIUS_PATH_LENGTH;
    p->trackPrefixStats = 0;
    return CX_OK;
}

int cx_renderer_create(cx_ctx* ctx, const cx_scene* scene, uint32_t w, uint32_t h, cx_renderer** out) {
    if (!ctx || !scene || !out) return set_err(ctx, CX_ERR_INVALID_ARG, "null argument");
    if (w == 0 || h == 0 || (uint64_t)w * h > (1ull << 28))
        return set_err(ctx, CX_ERR_INVALID_ARG, "image size must be in [1, 2^28] pixels");
    cx_renderer* r = new (std::nothrow) cx_renderer();
    if (!r) return set_err(ctx, CX_ERR_OUT_OF_MEMORY, "host allocation failed");
    r->ctx = ctx;
    r->scene = scene;
    r->W = w;
    r->H = h;
    r->P = w * h;
    cx_default_params(scene, w, h, &r->params);
    r->frameRngState = pcg_seed(0);
    hipError_t e = hipSuccess;
    auto A = [&](hipError_t x) {
        if (e == hipSuccess) e = x;
    };
    for (int b = 0; b < 2; b++) {
        A(dalloc(r->allocs, &r->prefixRes[b], (size_t)PREFIX_CHUNKS * r->P));
        A(dalloc(r->allocs, &r->suffixRes[b], (size_t)SUFFIX_CHUNKS * r->P));
    }
    A(dalloc(r->allocs, &r->gbuf, r->P));
    A(dalloc(r->allocs, &r->neighCount, r->P));
    A(dalloc(r->allocs, &r->counters, (size_t)COUNTER_STRIPES * NUM_COUNTERS));
    A(dalloc(r->allocs, &r->dyn, 1));
    A(dalloc(r->allocs, &r->devRng, 1));
    if (e != hipSuccess) {
        free_all(r->allocs);
        delete r;
        return hip_err(ctx, e, "renderer allocation");
    }
    int rc = entries_alloc(ctx, r->P, &r->entries);
    if (rc == CX_OK) rc = ensure_neighbors(r, 8);
    if (rc != CX_OK) {
        if (r->entries) cx_entries_destroy(r->entries);
        free_all(r->allocs);
        delete r;
        return rc;
    }
    rc = cx_renderer_reset(r, nullptr);
    if (rc == CX_OK) {
        e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) rc = hip_err(ctx, e, "renderer reset");
    }
    if (rc != CX_OK) {
        cx_renderer_destroy(r);
        return rc;
    }
    *out = r;
    return CX_OK;
}

int cx_renderer_destroy(cx_renderer* r) {
    if (!r) return CX_OK;
    if (r->side) (void)hipStreamSynchronize(r->side);
    for (hipEvent_t ev : {r->evFork, r->evQuery})
        if (ev) (void)hipEventDestroy(ev);
    if (r->side) (void)hipStreamDestroy(r->side);
    for (auto& g : r->graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (r->gstream) (void)hipStreamDestroy(r->gstream);
    if (r->evMade)
        for (auto& ev : r->ev) (void)hipEventDestroy(ev);
    for (auto& ev : r->traceEv) (void)hipEventDestroy(ev);
    if (r->entries) cx_entries_destroy(r->entries);
    free_all(r->wave.queueAllocs);
    free_all(r->wave.pixelAllocs);
    free_all(r->allocs);
    delete r;
    return CX_OK;
}

int cx_renderer_set_params(cx_renderer* r, const cx_launch_params* p) {
    if (!r || !p) return CX_ERR_INVALID_ARG;
    int rc = check_params(r, *p);
    if (rc != CX_OK) return rc;
    const uint32_t k = p->gatherM > 1 ? (uint32_t)(p->gatherM - 1) : 0u;
    rc = ensure_neighbors(r, std::max(k, 1u));
    if (rc != CX_OK) return rc;
    const uint32_t random = r->params.random;
    const uint64_t frameIdx = r->params.frameIdx;
    r->params = *p;
    r->params.random = random;
    r->params.frameIdx = frameIdx;
    r->params.width = r->W;
    r->params.height = r->H;
    return CX_OK;
}

int cx_renderer_get_params(const cx_renderer* r, cx_launch_params* p) {
    if (!r || !p) return CX_ERR_INVALID_ARG;
    *p = r->params;
    return CX_OK;
}

int cx_renderer_seed(cx_renderer* r, uint64_t seed) {
    if (!r) return CX_ERR_INVALID_ARG;
    r->frameRngState = pcg_seed(seed);
    return CX_OK;
}

int cx_renderer_reset(cx_renderer* r, void* stream) {
    if (!r) return CX_ERR_INVALID_ARG;
    r->front = 0;
    r->back = 0;
    r->frameIdx = 0;
    r->maxLenSeen = 0;
    r->maxPrefixSeen = 0;
    FrameArgs A = make_args(r);
    CX_HIP(r->ctx, launch_reset(A, (hipStream_t)stream));
    return CX_OK;
}

int cx_renderer_begin_frame(cx_renderer* r, const cx_camera* cam, int cameraChanged, float* out) {
    if (!r || !cam || !out) return set_err(r ? r->ctx : nullptr, CX_ERR_INVALID_ARG, "null argument");
    // src/Renderer.cpp:212-232
    ++r->frameIdx;
    if (cameraChanged || !r->params.enableAccum) r->frameIdx = 0;
    r->params.frameIdx = r->frameIdx;
    r->params.random = pcg_next(r->frameRngState);
    r->params.width = r->W;
    r->params.height = r->H;
    r->cam = *cam;
    r->out = out;
    if (r->front == 0) {
        r->front = 1;
        r->back = 0;
    } else {
        r->front = 0;
        r->back = 1;
    }
    r->params.spatialRoundIdx = 0;
    return CX_OK;
}

int cx_pass_launch(cx_renderer* r, cx_pass pass, const cx_launch_params* params, void* stream) {
    if (!r || !params) return CX_ERR_INVALID_ARG;
    if (!r->out) return set_err(r->ctx, CX_ERR_INVALID_ARG, "no frame begun (cx_renderer_begin_frame)");
    int prc = detect_caller_capture(r, stream);
    if (prc != CX_OK) return prc;
    prc = check_params(r, *params);
    if (prc != CX_OK) return prc;
    prc = ensure_neighbors(r, std::max(params->gatherM > 1 ? (uint32_t)(params->gatherM - 1) : 0u, 1u));
    if (prc != CX_OK) return prc;
    if (r->exec == CX_EXEC_WAVEFRONT) {
        prc = ensure_wave(r, *params);
        if (prc != CX_OK) return prc;
    }
    hipStream_t s = (hipStream_t)stream;
    const cx_launch_params saved = r->params;
    r->params = *params;
    r->params.width = r->W;
    r->params.height = r->H;
    FrameArgs A = make_args(r);
    int rc = CX_OK;
    hipError_t e = hipSuccess;
    switch (pass) {
        case CX_PASS_PREFIX_GEN_TEMP:
        case CX_PASS_PREFIX_SPATIAL:
        case CX_PASS_SUFFIX_GEN_TEMP:
        case CX_PASS_SUFFIX_SPATIAL:
            e = r->exec == CX_EXEC_WAVEFRONT ? wave_pass(r, pass, A, s) : launch_pass(pass, A, s);
            if (e == hipSuccess && pass == CX_PASS_PREFIX_SPATIAL)
                e = hipMemcpyAsync(r->prefixRes[r->back], r->prefixRes[r->front], (size_t)PREFIX_CHUNKS * r->P * 16,
                                   hipMemcpyDeviceToDevice, s);
            if (e == hipSuccess && pass == CX_PASS_SUFFIX_SPATIAL)
                e = hipMemcpyAsync(r->suffixRes[r->back], r->suffixRes[r->front], (size_t)SUFFIX_CHUNKS * r->P * 16,
                                   hipMemcpyDeviceToDevice, s);
            break;
        case CX_PASS_PREFIX_STORE_ENTRIES: {
            const float* mn = r->scene->info.boundsMin;
            const float* mx = r->scene->info.boundsMax;
            set_quantisation(r->entries->E, mn, mx);
            e = launch_store_entries(A, r->entries->E, s);
            break;
        }
        case CX_PASS_ENTRIES_BUILD:
            e = launch_tree_build(r->entries->E, r->entries->T, r->P, r->entries->sortTemp, r->entries->sortTempBytes, s);
            break;
        case CX_PASS_FINAL_GATHER:
            e = r->exec == CX_EXEC_WAVEFRONT ? wave_pass(r, pass, A, s) : launch_pass(pass, A, s);
            break;
        default: rc = set_err(r->ctx, CX_ERR_INVALID_ARG, "unknown pass");
    }
    r->params = saved;
    r->maxLenSeen = std::max(r->maxLenSeen, (uint32_t)params->maxPathLen);
    r->maxPrefixSeen = std::max(r->maxPrefixSeen, prefix_max_len(*params));
    if (e != hipSuccess) return hip_err(r->ctx, e, "pass launch");
    return rc;
}

}  // extern "C"

namespace {

// The side-stream fork of the final phase: not inside a caller's stream capture, where the
// frame keeps to the caller's stream (the recorded graph is then a chain of kernel and copy
// nodes, the form the caller-capture tests cover).
bool overlap_final(const cx_renderer* r) {
    return r->overlap && !r->callerCapture && r->exec == CX_EXEC_WAVEFRONT &&
           r->params.rendererType == CX_CONDITIONAL_RESTIR;
}

int ensure_side(cx_renderer* r) {
    if (r->side) return CX_OK;
    CX_HIP(r->ctx, hipStreamCreateWithFlags(&r->side, hipStreamNonBlocking));
    for (hipEvent_t* ev : {&r->evFork, &r->evQuery})
        CX_HIP(r->ctx, hipEventCreateWithFlags(ev, hipEventDisableTiming));
    return CX_OK;
}

// The launches of one frame after cx_renderer_begin_frame (src/Renderer.cpp:259-371); also
// what a graph-mode capture records.
int frame_body(cx_renderer* r, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    cx_launch_params& lp = r->params;
    int rc = CX_OK;
    for (bool& b : r->passRan) b = false;
    r->traceEvUsed = 0;
    r->traceEvOk = true;
    if (r->deviceRandom) {   // camera and frame index from the host, the random number on the device
        FrameDyn d;
        std::memset(&d, 0, sizeof(d));
        std::memcpy(d.cam, r->cam.pos, 12);
        std::memcpy(d.cam + 3, r->cam.U, 12);
        std::memcpy(d.cam + 6, r->cam.V, 12);
        std::memcpy(d.cam + 9, r->cam.W, 12);
        d.frameIdx = r->params.frameIdx;
        CX_HIP(r->ctx, launch_next_dyn(r->dyn, r->devRng, d, s));
    }
    CX_HIP(r->ctx, launch_zero_counters(r->counters, COUNTER_STRIPES * NUM_COUNTERS, s));
    rec(r, 0, s);
    auto run = [&](cx_pass p) -> int {
        r->passRan[p] = true;
        return cx_pass_launch(r, p, &lp, stream);
    };
    if ((rc = run(CX_PASS_PREFIX_GEN_TEMP))) return rc;
    rec(r, 1, s);
    const bool restir = lp.rendererType == CX_CONDITIONAL_RESTIR || lp.rendererType == CX_RESTIR_PT;
    if (restir && lp.prefixEnableSpatial) {
        for (int i = 0; i < lp.prefixSpatialRounds; i++) {
            lp.spatialRoundIdx = i;
            if ((rc = run(CX_PASS_PREFIX_SPATIAL))) return rc;
        }
    }
    rec(r, 2, s);
    const bool cond = lp.rendererType == CX_CONDITIONAL_RESTIR;
    if (cond) {
        if ((rc = run(CX_PASS_SUFFIX_GEN_TEMP))) return rc;
    }
    rec(r, 3, s);
    if (cond && lp.suffixEnableSpatial && lp.suffixSpatialCount > 0) {
        for (int i = 0; i < lp.suffixSpatialRounds; i++) {
            lp.spatialRoundIdx = i;
            if ((rc = run(CX_PASS_SUFFIX_SPATIAL))) return rc;
        }
    }
    rec(r, 4, s);
    // With the overlap, the entry store and build go to the side stream, forked here, and the
    // final gather joins it before its shifts; their time is then inside the final gather's.
    const bool overlap = cond && overlap_final(r);
    if (overlap) {
        if ((rc = ensure_side(r))) return rc;
        CX_HIP(r->ctx, hipEventRecord(r->evFork, s));
        CX_HIP(r->ctx, hipStreamWaitEvent(r->side, r->evFork, 0));
    }
    void* entryStream = overlap ? (void*)r->side : stream;
    auto runOn = [&](cx_pass p, void* st) -> int {
        r->passRan[p] = true;
        return cx_pass_launch(r, p, &lp, st);
    };
    if (cond) {
        if ((rc = runOn(CX_PASS_PREFIX_STORE_ENTRIES, entryStream))) return rc;
    }
    rec(r, 5, s);
    if (cond) {
        if ((rc = runOn(CX_PASS_ENTRIES_BUILD, entryStream))) return rc;
    }
    rec(r, 6, s);
    if (cond) {
        r->fgSide = overlap ? r->side : nullptr;
        rc = run(CX_PASS_FINAL_GATHER);
        r->fgSide = nullptr;
        if (rc) return rc;
    }
    rec(r, 7, s);
    r->evValid = r->timing;
    return CX_OK;
}


// What a captured frame bakes in besides its parity: every parameter except the per-frame
// random and index, the output buffer, the execution settings, the round bounds and the
// buffers' allocation epoch.
std::vector<uint8_t> graph_key(const cx_renderer* r) {
    cx_launch_params p = r->params;
    p.random = 0;
    p.frameIdx = 0;
    p.spatialRoundIdx = 0;
    std::vector<uint8_t> k(sizeof(p));
    std::memcpy(k.data(), &p, sizeof(p));
    auto put = [&](const void* v, size_t n) { k.insert(k.end(), (const uint8_t*)v, (const uint8_t*)v + n); };
    put(&r->out, sizeof(r->out));
    put(&r->exec, sizeof(r->exec));
    put(&r->gridLimit, sizeof(r->gridLimit));
    put(&r->maxLenSeen, sizeof(r->maxLenSeen));
    put(&r->maxPrefixSeen, sizeof(r->maxPrefixSeen));
    put(&r->allocEpoch, sizeof(r->allocEpoch));
    put(&r->overlap, sizeof(r->overlap));
    put(&r->waveRounds, sizeof(r->waveRounds));
    put(&r->deviceRandom, sizeof(r->deviceRandom));
    return k;
}

int graph_frame(cx_renderer* r, void* stream) {
    hipStream_t s = (hipStream_t)stream;
    if (!s) {
        // the legacy null stream cannot be captured; a blocking stream is ordered with it
        if (!r->gstream) CX_HIP(r->ctx, hipStreamCreate(&r->gstream));
        s = r->gstream;
    }
    // every buffer the frame uses exists before the capture (no allocation inside it)
    const cx_launch_params& lp = r->params;
    int rc = ensure_neighbors(r, std::max(lp.gatherM > 1 ? (uint32_t)(lp.gatherM - 1) : 0u, 1u));
    if (rc == CX_OK && r->exec == CX_EXEC_WAVEFRONT) rc = ensure_wave(r, lp);
    if (rc == CX_OK && overlap_final(r)) rc = ensure_side(r);
    if (rc != CX_OK) return rc;
    auto& g = r->graphs[r->front];
    if (!g.exec || g.key != graph_key(r)) {
        if (g.exec) {
            CX_HIP(r->ctx, hipStreamSynchronize(s));
            (void)hipGraphExecDestroy(g.exec);
            g.exec = nullptr;
        }
        CX_HIP(r->ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        r->capturing = true;
        rc = frame_body(r, s);
        r->capturing = false;
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &graph);
        if (rc != CX_OK || e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc != CX_OK ? rc : hip_err(r->ctx, e, "frame capture");
        }
        const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            g.exec = nullptr;
            return hip_err(r->ctx, ei, "frame graph instantiation");
        }
        // the rounds the capture used are max(seen, current), which is what `seen` now holds
        g.key = graph_key(r);
        r->captures++;
    }
    FrameDyn d;
    std::memset(&d, 0, sizeof(d));
    std::memcpy(d.cam, r->cam.pos, 12);
    std::memcpy(d.cam + 3, r->cam.U, 12);
    std::memcpy(d.cam + 6, r->cam.V, 12);
    std::memcpy(d.cam + 9, r->cam.W, 12);
    d.random = r->params.random;
    d.frameIdx = r->params.frameIdx;
    CX_HIP(r->ctx, launch_set_dyn(r->dyn, d, s));
    CX_HIP(r->ctx, hipGraphLaunch(g.exec, s));
    return CX_OK;
}

}  // namespace

extern "C" {

int cx_renderer_frame(cx_renderer* r, const cx_camera* cam, int cameraChanged, float* out, void* stream) {
    if (!r) return CX_ERR_INVALID_ARG;
    int rc = detect_caller_capture(r, stream);
    if (rc != CX_OK) return rc;
    if (r->callerCapture) {
        if (r->graphMode || r->timing)
            return set_err(r->ctx, CX_ERR_INVALID_ARG,
                           "recording a frame into the caller's graph needs graph mode and timing off (they "
                           "record or create their own graphs and events)");
        // every buffer must exist already: these only check, and refuse growth, under capture
        const cx_launch_params& lp = r->params;
        rc = ensure_neighbors(r, std::max(lp.gatherM > 1 ? (uint32_t)(lp.gatherM - 1) : 0u, 1u));
        if (rc == CX_OK && r->exec == CX_EXEC_WAVEFRONT) rc = ensure_wave(r, lp);
        if (rc != CX_OK) return rc;
    }
    rc = cx_renderer_begin_frame(r, cam, cameraChanged, out);
    if (rc != CX_OK) return rc;
    if (r->graphMode && !r->timing) {
        r->evValid = false;
        return graph_frame(r, stream);
    }
    if (r->timing && !r->evMade) {
        // Timing-only events: no system-scope fence (cache writeback and invalidation) at
        // each record, which would both slow the frame and inflate the measured intervals.
        for (auto& ev : r->ev) CX_HIP(r->ctx, hipEventCreateWithFlags(&ev, hipEventDisableSystemFence));
        r->evMade = true;
    }
    return frame_body(r, stream);
}

int cx_renderer_set_graph_mode(cx_renderer* r, int enable) {
    if (!r) return CX_ERR_INVALID_ARG;
    r->graphMode = enable != 0;
    return CX_OK;
}

int cx_renderer_graph_captures(const cx_renderer* r) { return r ? (int)r->captures : -1; }

int cx_renderer_set_wave_rounds(cx_renderer* r, uint32_t rounds) {
    if (!r) return CX_ERR_INVALID_ARG;
    r->waveRounds = rounds;
    return CX_OK;
}

int cx_renderer_set_device_random(cx_renderer* r, int enable, uint64_t seed) {
    if (!r) return CX_ERR_INVALID_ARG;
    if (enable) {
        const uint64_t st = pcg_seed(seed);
        CX_HIP(r->ctx, hipMemcpy(r->devRng, &st, sizeof(st), hipMemcpyHostToDevice));
    }
    r->deviceRandom = enable != 0;
    return CX_OK;
}

int cx_scene_set_materials_device(cx_scene* s, const float* rows, uint32_t n, int tables, void* stream) {
    if (!s || (n && !rows)) return CX_ERR_INVALID_ARG;
    if (n != s->numMaterials) return set_err(s->ctx, CX_ERR_INVALID_ARG, "material count differs from the scene's");
    if (tables & ~(CX_TABLE_SAMPLING | CX_TABLE_EVAL)) return set_err(s->ctx, CX_ERR_INVALID_ARG, "unknown table");
    const hipStream_t st = (hipStream_t)stream;
    if (tables & CX_TABLE_SAMPLING) CX_HIP(s->ctx, launch_set_materials((DMaterial*)s->d.mats, rows, n, st));
    if (tables & CX_TABLE_EVAL) CX_HIP(s->ctx, launch_set_materials(s->evalMats, rows, n, st));
    return CX_OK;
}

int cx_renderer_set_overlap(cx_renderer* r, int enable) {
    if (!r) return CX_ERR_INVALID_ARG;
    r->overlap = enable != 0;
    return CX_OK;
}

int cx_renderer_set_queue_grid(cx_renderer* r, uint32_t maxBlocks) {
    if (!r) return CX_ERR_INVALID_ARG;
    if (maxBlocks < 1 || maxBlocks > (1u << 20)) return set_err(r->ctx, CX_ERR_INVALID_ARG, "queue grid out of range");
    r->gridLimit = maxBlocks;
    return CX_OK;
}

int cx_batch_frame(cx_ctx* ctx, cx_renderer* const* rs, const cx_camera* cams, const int* cameraChanged,
                   float* const* outs, uint32_t n, void* const* streams) {
    if (!ctx || (n && (!rs || !cams || !cameraChanged || !outs))) return set_err(ctx, CX_ERR_INVALID_ARG, "null argument");
    for (uint32_t i = 0; i < n; i++) {
        if (!rs[i] || !outs[i]) return set_err(ctx, CX_ERR_INVALID_ARG, "null renderer or output");
        for (uint32_t j = 0; j < i; j++)
            if (rs[j] == rs[i]) return set_err(ctx, CX_ERR_INVALID_ARG, "a renderer appears twice in the batch");
    }
    while (!streams && ctx->pool.size() < n) {
        hipStream_t s = nullptr;
        CX_HIP(ctx, hipStreamCreate(&s));
        ctx->pool.push_back(s);
    }
    for (uint32_t i = 0; i < n; i++) {
        void* s = streams ? streams[i] : (void*)ctx->pool[i];
        rs[i]->concurrent = n > 1;
        const int rc = cx_renderer_frame(rs[i], &cams[i], cameraChanged[i], outs[i], s);
        rs[i]->concurrent = false;
        if (rc != CX_OK) {
            if (rs[i]->ctx != ctx) set_err(ctx, rc, cx_last_error(rs[i]->ctx));
            return rc;
        }
    }
    return CX_OK;
}

int cx_batch_sync(cx_ctx* ctx) {
    if (!ctx) return CX_ERR_INVALID_ARG;
    for (hipStream_t s : ctx->pool) CX_HIP(ctx, hipStreamSynchronize(s));
    return CX_OK;
}

int cx_renderer_set_exec_mode(cx_renderer* r, int mode) {
    if (!r) return CX_ERR_INVALID_ARG;
    if (mode != CX_EXEC_WAVEFRONT && mode != CX_EXEC_MEGAKERNEL)
        return set_err(r->ctx, CX_ERR_INVALID_ARG, "unknown execution mode");
    r->exec = mode;
    return CX_OK;
}

int cx_renderer_get_exec_mode(const cx_renderer* r) { return r ? r->exec : -1; }

int cx_renderer_enable_timing(cx_renderer* r, int enable) {
    if (!r) return CX_ERR_INVALID_ARG;
    r->timing = enable != 0;
    return CX_OK;
}

int cx_renderer_get_stats(cx_renderer* r, cx_frame_stats* st) {
    if (!r || !st) return CX_ERR_INVALID_ARG;
    std::memset(st, 0, sizeof(*st));
    std::vector<unsigned long long> cs((size_t)COUNTER_STRIPES * NUM_COUNTERS);
    CX_HIP(r->ctx, hipDeviceSynchronize());
    CX_HIP(r->ctx, hipMemcpy(cs.data(), r->counters, cs.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    unsigned long long c[NUM_COUNTERS] = {0};
    c[STAT_MIN] = ~0ull;
    for (int stp = 0; stp < COUNTER_STRIPES; stp++)
        for (int i = 0; i < NUM_COUNTERS; i++) {
            const unsigned long long x = cs[(size_t)stp * NUM_COUNTERS + i];
            if (i == STAT_MIN) c[i] = std::min(c[i], x);
            else if (i == STAT_MAX) c[i] = std::max(c[i], x);
            else c[i] += x;
        }
    for (int i = 0; i < CX_PASS_COUNT; i++) {
        st->passClosestRays[i] = c[2 * i];
        st->passShadowRays[i] = c[2 * i + 1];
        st->closestRays += c[2 * i];
        st->shadowRays += c[2 * i + 1];
    }
    st->tailClosestRays = c[STAT_TAIL_CLOSEST];
    st->tailShadowRays = c[STAT_TAIL_SHADOW];
    st->totalNeighCount = c[STAT_TOTAL];
    st->minNeighCount = (uint32_t)std::min<unsigned long long>(c[STAT_MIN], 0xFFFFFFFFull);
    st->maxNeighCount = (uint32_t)c[STAT_MAX];
    uint32_t nv = 0;
    CX_HIP(r->ctx, hipMemcpy(&nv, r->entries->E.numValid, 4, hipMemcpyDeviceToHost));
    st->numEntries = nv;
    if (r->wave.count) {
        uint32_t ovf = 0;
        CX_HIP(r->ctx, hipMemcpy(&ovf, r->wave.count + (size_t)2 * QSTRIPES * QCOUNT_STRIDE, 4, hipMemcpyDeviceToHost));
        st->queueOverflow = ovf;
    }
    if (r->evValid) {
        for (int i = 0; i < CX_PASS_COUNT; i++) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, r->ev[i], r->ev[i + 1]);
            st->passMs[i] = ms;
        }
        (void)hipEventElapsedTime(&st->totalMs, r->ev[0], r->ev[CX_PASS_COUNT]);
        if (r->traceEvOk) {
            st->traceLaunches = r->traceEvUsed / 2;
            for (uint32_t i = 0; i + 1 < r->traceEvUsed; i += 2) {
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, r->traceEv[i], r->traceEv[i + 1]);
                st->traceMs += ms;
            }
        }
    }
    return CX_OK;
}

int cx_renderer_adjoint(cx_renderer* r, const float* dLdI, double* gAlbedo, double* gRough, double* gEmit,
                        float* replay, uint32_t* status, void* stream) {
    if (!r) return CX_ERR_INVALID_ARG;
    const int type = r->params.rendererType;
    if (type != CX_RESTIR_PT && type != CX_PATH_TRACER)
        return set_err(r->ctx, CX_ERR_INVALID_ARG, "the adjoint covers ReSTIR PT and the path tracer");
    if (!r->out) return set_err(r->ctx, CX_ERR_INVALID_ARG, "no frame rendered yet");
    if ((gAlbedo || gRough || gEmit) && !dLdI) return set_err(r->ctx, CX_ERR_INVALID_ARG, "gradients need dLdI");
    FrameArgs A = make_args(r);
    AdjArgs G;
    G.dLdI = dLdI;
    G.gAlbedo = gAlbedo;
    G.gRough = gRough;
    G.gEmit = gEmit;
    G.replay = replay;
    G.status = status;
    G.evalMats = r->scene->evalMats;
    G.evalEmit = r->scene->evalEmit;
    G.numMaterials = r->scene->numMaterials;
    G.numEmitters = r->scene->d.numEmitters;
    G.ldsSlots = 0;
    CX_HIP(r->ctx, launch_adjoint(A, G, (hipStream_t)stream));
    return CX_OK;
}

int cx_renderer_state(cx_renderer* r, cx_state_view* v) {
    if (!r || !v) return CX_ERR_INVALID_ARG;
    std::memset(v, 0, sizeof(*v));
    for (int b = 0; b < 2; b++) {
        v->prefixRes[b] = r->prefixRes[b];
        v->suffixRes[b] = r->suffixRes[b];
    }
    v->gbuffer = r->gbuf;
    v->neighIdx = r->neighIdx;
    v->neighDist = r->neighDist;
    v->neighCount = r->neighCount;
    v->entryPos = r->entries->E.pos;
    v->entryValid = r->entries->E.valid;
    v->frontBufferIdx = r->front;
    v->backBufferIdx = r->back;
    v->prefixChunks = PREFIX_CHUNKS;
    v->suffixChunks = SUFFIX_CHUNKS;
    v->k = r->params.gatherM > 1 ? (uint32_t)(r->params.gatherM - 1) : 0u;
    return CX_OK;
}

}  // extern "C"
