// context.cpp — the C ABI (include/astraea_hip.h): device memory, layout, field transfer,
// stage calls, the fused KDK step with hipGraph replay, stage timing and reductions.
//
// The reference keeps this in Simulation::Impl (engine/simulation.cpp:345-536) and the launch
// wrappers of engine/sph_gravity.cu:1986-2161; here one context owns one HIP stream and all
// device memory, and the step never waits on the host (dt stays on the device).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/astraea_hip.h"
#include "errors.hpp"
#include "launch.hpp"

using namespace astraea_hip;

namespace {

thread_local std::string gLastError;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    gLastError = buf;
    return code;
}

#define HIP_TRY(expr)                                                                                  \
    do {                                                                                               \
        const hipError_t e__ = (expr);                                                                 \
        if (e__ != hipSuccess) return fail(ASTRAEA_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e__)); \
    } while (0)

uint32_t roundUp(uint32_t v, uint32_t a) { return (v + a - 1) / a * a; }

enum Kind { kF32, kU8, kU32, kF64 };

struct FieldLoc {
    int kind;        // element type on the host
    int set;         // 0 persistent (cur), 1 derived
    void** arr;      // address of the device pointer
    int comps;       // components per device element (4 = float4, 2 = float2, 1 = scalar)
    int comp;        // component index (for u8 inside meta: byte index)
};

// Table of the allocation-free render entry point (astraea_pack_render_noalloc): for each of at
// most ASTRAEA_NOALLOC_MAX_SEGMENTS segments, the centre (3 floats, the first 3 x max entries) and
// the output offset (1 u32, the last max entries).  4 KiB per GPU, held until the process ends.
// Host code never uses this name as an address: each device's copy is reached through the
// address hipGetSymbolAddress returns with that device current (scratchAddress).
constexpr uint32_t kScratchWords = 4u * ASTRAEA_NOALLOC_MAX_SEGMENTS;
static __device__ __align__(256) uint32_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
uint32_t* gScratchAddr[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be the current device; looked up once
// per device.
int scratchAddress(int device, uint32_t** out) {
    if (device < 0 || device >= kMaxDevices) return fail(ASTRAEA_E_INVALID, "device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (!gScratchAddr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(ASTRAEA_E_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", device, hipGetErrorString(e));
        gScratchAddr[device] = static_cast<uint32_t*>(p);
    }
    *out = gScratchAddr[device];
    return ASTRAEA_OK;
}

}  // namespace

int astraea_hip::reportError(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    gLastError = buf;
    return code;
}

struct astraea_ctx {
    int device = 0;
    uint32_t* scratchDev = nullptr;        // this device's address of the file-scope `scratch` table
    hipStream_t stream = nullptr;
    // gravity runs on `side`, beside the neighbour build and pass 1 on `stream` (it reads only the
    // tree and the positions; pass 2 waits for it): fork and join events
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    uint32_t capSlots = 0;
    uint32_t maxSeg = 0;
    astraea_params params{};
    std::vector<uint32_t> counts, offs;
    uint32_t numSeg = 0, slots = 0;
    DeviceCtx dc{};
    std::vector<void*> allocs;
    double* dEnergy = nullptr;
    double* dClass = nullptr;              // ASTRAEA_CLASS_STRIDE per segment
    double* dRem = nullptr;                // 3 x 8 per segment: remnant guess | sums | body 0
    unsigned long long* dSeed = nullptr;   // remnant seed key per segment
    float* dAnchor = nullptr;              // relaxation anchor temperatures by original id (capSlots)
    // render frames
    uint32_t* dRender = nullptr;           // 5 dwords per particle (capSlots)
    uint8_t* dSpeed = nullptr;             // capSlots
    float* dCom = nullptr;                 // 2 x 3 floats per segment
    uint32_t* dOutBase = nullptr;          // per segment
    // crust tracers: tables by tracer and by original id, never permuted
    int32_t* dTracerOf = nullptr;          // tracer of each original id, -1 for none (capSlots)
    float4* dTracerRest = nullptr;         // (rest radius, rest normal) per tracer
    uint8_t* dTracerOcean = nullptr;
    uint32_t* dTracerOut = nullptr;        // 10 dwords per tracer
    uint32_t tracerCount = 0, tracerCap = 0;
    bool loaded = false, treeValid = false, listValid = false, dtRawValid = false, posNewPending = false;
    int parity = 0;  // number of cur/alt swaps mod 2
    // steps queued on the stream: kStepEnd launches since the layout (the device counts the ones
    // that completed, Counters kCntSteps), and the dt_max of every astraea_step step not yet
    // settled, so that steps held by an incomplete neighbour table can be run again
    uint64_t stepsQueued = 0;
    std::vector<float> unsettled;
    uint64_t stepsRedone = 0;      // held steps finished or run again after a table was widened
    // graph
    hipGraphExec_t graph = nullptr;
    float graphMaxDt = -1.0f;
    int graphParity = -1;
    // timing: event pairs recorded on the context stream, read back lazily
    bool timing = false;
    static constexpr int kPool = 1024;
    std::vector<hipEvent_t> evPool;
    struct Rec { int stage, a, b; };
    std::vector<Rec> pending;
    int evNext = 0;
    int openEv[ASTRAEA_STAGE_COUNT_] = {};
    double stageMs[ASTRAEA_STAGE_COUNT_] = {};
    uint64_t stageN[ASTRAEA_STAGE_COUNT_] = {};
};

namespace {

template <class T>
int devAlloc(astraea_ctx* c, T*& p, size_t count) {
    void* q = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    const hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return fail(ASTRAEA_E_HIP, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    c->allocs.push_back(q);
    p = static_cast<T*>(q);
    return ASTRAEA_OK;
}

#define TRY(x)                      \
    do {                            \
        const int r__ = (x);        \
        if (r__ != ASTRAEA_OK) return r__; \
    } while (0)

int allocPersist(astraea_ctx* c, Persist& p, uint32_t cap) {
    TRY(devAlloc(c, p.pos, cap));
    TRY(devAlloc(c, p.vel, cap));
    TRY(devAlloc(c, p.hu, cap));
    TRY(devAlloc(c, p.str0, cap));
    TRY(devAlloc(c, p.str1, cap));
    TRY(devAlloc(c, p.meta, cap));
    TRY(devAlloc(c, p.oid, cap));
    TRY(devAlloc(c, p.p64, 3 * (size_t)cap));
    TRY(devAlloc(c, p.v64, 3 * (size_t)cap));
    return ASTRAEA_OK;
}

void freeGraph(astraea_ctx* c) {
    if (c->graph) {
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    c->graphParity = -1;
}

constexpr uint32_t kMaxNeighbourStride = 1u << 16;

// Union entries a wavefront may hold for a table of `stride` entries per slot: as many bytes as
// the entries themselves, and at most 2^16 (the entries are 16-bit positions).
uint32_t unionCapOf(uint32_t stride) { return (uint32_t)std::min<uint64_t>(32ull * stride, 1ull << 16); }

template <class T>
void releaseArray(astraea_ctx* c, T*& p) {
    if (!p) return;
    (void)hipFree(p);
    auto it = std::find(c->allocs.begin(), c->allocs.end(), (void*)p);
    if (it != c->allocs.end()) c->allocs.erase(it);
    p = nullptr;
}

int allocNeighbourList(astraea_ctx* c, uint32_t stride) {
    releaseArray(c, c->dc.nb.list);
    releaseArray(c, c->dc.nb.unionIds);
    const uint32_t uc = unionCapOf(stride);
    TRY(devAlloc(c, c->dc.nb.list, (size_t)stride * c->capSlots));
    TRY(devAlloc(c, c->dc.nb.unionIds, (size_t)uc * (c->capSlots / kWave)));
    c->dc.nb.stride = stride;
    c->dc.nb.unionCap = uc;
    freeGraph(c);
    return ASTRAEA_OK;
}

// The spill pool of the wavefronts the regular region cannot hold (a captured step holds its
// address, so the graph goes with it).
int allocSpillPool(astraea_ctx* c, unsigned long long bytes) {
    releaseArray(c, c->dc.nb.pool);
    bytes = std::max<unsigned long long>((bytes + 255ull) & ~255ull, 256ull);
    TRY(devAlloc(c, c->dc.nb.pool, (size_t)bytes));
    c->dc.nb.poolBytes = bytes;
    freeGraph(c);
    return ASTRAEA_OK;
}

Params toDeviceParams(const astraea_params& p) {
    Params d;
    const float th = std::max(0.0f, p.theta);
    d.theta2 = th * th;
    {
        int e = 0;
        d.theta2Pow2 = d.theta2 > 0.0f && std::isfinite(d.theta2) && std::frexp(d.theta2, &e) == 0.5f &&
                       std::isnormal(d.theta2) ? 1u : 0u;
    }
    d.softFactor = p.soft_factor;
    d.cfl = p.cfl_factor;
    d.gammaHHe = p.gamma_hhe;
    d.coolMinT = p.cooling_min_temp_k;
    d.strengthMask = p.strength_mask & 0x1fu;
    d.group = p.group_size == 32u ? 32u : 64u;
    return d;
}

void syncDerivedFlags(astraea_ctx* c) {
    c->dc.prm = toDeviceParams(c->params);
    c->dc.fp64 = c->params.precision == ASTRAEA_PREC_FP64_ACCUM;
    c->dc.strength = c->dc.prm.strengthMask != 0u;
}

bool locate(astraea_ctx* c, int f, FieldLoc& L) {
    Persist& p = c->dc.cur;
    Derived& d = c->dc.d;
    auto F4 = [&](float4*& a, int comp, int set) { L = {kF32, set, (void**)&a, 4, comp}; };
    switch (f) {
        case ASTRAEA_F_POS_X: case ASTRAEA_F_POS_Y: case ASTRAEA_F_POS_Z: F4(p.pos, f - ASTRAEA_F_POS_X, 0); return true;
        case ASTRAEA_F_MASS: F4(p.pos, 3, 0); return true;
        case ASTRAEA_F_VEL_X: case ASTRAEA_F_VEL_Y: case ASTRAEA_F_VEL_Z: F4(p.vel, f - ASTRAEA_F_VEL_X, 0); return true;
        case ASTRAEA_F_RHO: F4(p.vel, 3, 0); return true;
        case ASTRAEA_F_H: F4(p.hu, 0, 0); return true;
        case ASTRAEA_F_H0: F4(p.hu, 1, 0); return true;
        case ASTRAEA_F_U: F4(p.hu, 2, 0); return true;
        case ASTRAEA_F_DAMAGE: F4(p.hu, 3, 0); return true;
        case ASTRAEA_F_SXX: case ASTRAEA_F_SYY: case ASTRAEA_F_SZZ: case ASTRAEA_F_SXY: F4(p.str0, f - ASTRAEA_F_SXX, 0); return true;
        case ASTRAEA_F_SXZ: F4(p.str1, 0, 0); return true;
        case ASTRAEA_F_SYZ: F4(p.str1, 1, 0); return true;
        case ASTRAEA_F_DAMAGE_CBRT: F4(p.str1, 2, 0); return true;
        case ASTRAEA_F_EPS_FAIL: F4(p.str1, 3, 0); return true;
        case ASTRAEA_F_MATERIAL: L = {kU8, 0, (void**)&p.meta, 1, 0}; return true;
        case ASTRAEA_F_BODY: L = {kU8, 0, (void**)&p.meta, 1, 1}; return true;
        case ASTRAEA_F_TRACER: L = {kU8, 0, (void**)&p.meta, 1, 2}; return true;
        case ASTRAEA_F_ORIGINAL_ID: L = {kU32, 0, (void**)&p.oid, 1, 0}; return true;
        case ASTRAEA_F_POS64_X: case ASTRAEA_F_POS64_Y: case ASTRAEA_F_POS64_Z:
            L = {kF64, 0, (void**)&p.p64, 1, f - ASTRAEA_F_POS64_X}; return true;
        case ASTRAEA_F_VEL64_X: case ASTRAEA_F_VEL64_Y: case ASTRAEA_F_VEL64_Z:
            L = {kF64, 0, (void**)&p.v64, 1, f - ASTRAEA_F_VEL64_X}; return true;
        case ASTRAEA_F_ACC_X: case ASTRAEA_F_ACC_Y: case ASTRAEA_F_ACC_Z: F4(d.rate, f - ASTRAEA_F_ACC_X, 1); return true;
        case ASTRAEA_F_DRHO_DT: F4(d.rate, 3, 1); return true;
        case ASTRAEA_F_ACC_GRAV_X: case ASTRAEA_F_ACC_GRAV_Y: case ASTRAEA_F_ACC_GRAV_Z:
            F4(d.grav, f - ASTRAEA_F_ACC_GRAV_X, 1); return true;
        case ASTRAEA_F_PHI: F4(d.grav, 3, 1); return true;
        case ASTRAEA_F_DU_DT: F4(d.rate2, 0, 1); return true;
        case ASTRAEA_F_DDAMAGE_CBRT_DT: F4(d.rate2, 1, 1); return true;
        case ASTRAEA_F_RAD_POWER_W: F4(d.rate2, 2, 1); return true;
        case ASTRAEA_F_PRESSURE: F4(d.eos, 0, 1); return true;
        case ASTRAEA_F_SOUND_SPEED: F4(d.eos, 1, 1); return true;
        case ASTRAEA_F_TEMP_K: F4(d.eos, 2, 1); return true;
        case ASTRAEA_F_SHEPARD: F4(d.p1a, 0, 1); return true;
        case ASTRAEA_F_SURF_FACTOR: F4(d.p1a, 1, 1); return true;
        case ASTRAEA_F_DIV_V: F4(d.p1a, 2, 1); return true;
        case ASTRAEA_F_NGB_COUNT: L = {kU32, 1, (void**)&d.p1a, 4, 3}; return true;
        case ASTRAEA_F_CURL_X: case ASTRAEA_F_CURL_Y: case ASTRAEA_F_CURL_Z: F4(d.p1b, f - ASTRAEA_F_CURL_X, 1); return true;
        case ASTRAEA_F_DS_XX: case ASTRAEA_F_DS_YY: case ASTRAEA_F_DS_ZZ: case ASTRAEA_F_DS_XY:
            F4(d.dS0, f - ASTRAEA_F_DS_XX, 1); return true;
        case ASTRAEA_F_DS_XZ: L = {kF32, 1, (void**)&d.dS1, 2, 0}; return true;
        case ASTRAEA_F_DS_YZ: L = {kF32, 1, (void**)&d.dS1, 2, 1}; return true;
        case ASTRAEA_F_PHASE: L = {kU8, 1, (void**)&d.phase, 0, 0}; return true;
        default: return false;
    }
}

// Device element bytes of an array described by a FieldLoc.
size_t elemBytes(const FieldLoc& L) {
    if (L.kind == kF64) return sizeof(double);
    if (L.arr == nullptr) return 0;
    if (L.kind == kU8 && L.comps == 0) return 1;   // phase: plain u8 array
    if (L.kind == kU8) return 4;                   // meta: u32
    if (L.kind == kU32 && L.comps == 1) return 4;
    return (size_t)L.comps * 4;
}

// Copy a whole padded device array to the host (or back).
int copyArray(astraea_ctx* c, void* dev, std::vector<uint8_t>& host, size_t bytesPerElem, bool toHost) {
    const size_t bytes = bytesPerElem * (size_t)c->slots;
    if (toHost) {
        host.resize(bytes);
        HIP_TRY(hipMemcpyAsync(host.data(), dev, bytes, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    } else {
        HIP_TRY(hipMemcpyAsync(dev, host.data(), bytes, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return ASTRAEA_OK;
}

// Run `fn(hostIndex, slot)` over every particle.
template <class F>
void forEachParticle(const astraea_ctx* c, F&& fn) {
    size_t h = 0;
    for (uint32_t s = 0; s < c->numSeg; ++s)
        for (uint32_t k = 0; k < c->counts[s]; ++k) fn(h++, (size_t)c->offs[s] + k);
}

int transferField(astraea_ctx* c, int f, void* hostBuf, bool upload) {
    FieldLoc L;
    if (!locate(c, f, L)) return fail(ASTRAEA_E_INVALID, "unknown field id %d", f);
    if (upload && L.set != 0) return fail(ASTRAEA_E_INVALID, "field %d is derived and cannot be uploaded", f);
    if (!hostBuf) return fail(ASTRAEA_E_INVALID, "null host buffer");
    // A drift leaves the new positions in posNew until the next sort.
    const bool posPending = !upload && c->posNewPending && (f >= ASTRAEA_F_POS_X && f <= ASTRAEA_F_POS_Z);
    if (L.kind == kF64) {
        double* base = (double*)(*L.arr) + (size_t)L.comp * c->capSlots;
        std::vector<uint8_t> buf;
        TRY(copyArray(c, base, buf, sizeof(double), true));
        double* d = (double*)buf.data();
        if (upload) {
            forEachParticle(c, [&](size_t h, size_t s) { d[s] = ((const double*)hostBuf)[h]; });
            TRY(copyArray(c, base, buf, sizeof(double), false));
            // keep the FP32 copy in sync
            const int f32 = (f <= ASTRAEA_F_POS64_Z) ? ASTRAEA_F_POS_X + (f - ASTRAEA_F_POS64_X) : ASTRAEA_F_VEL_X + (f - ASTRAEA_F_VEL64_X);
            std::vector<float> hf;
            size_t n = 0;
            for (uint32_t s = 0; s < c->numSeg; ++s) n += c->counts[s];
            hf.resize(n);
            for (size_t k = 0; k < n; ++k) hf[k] = (float)((const double*)hostBuf)[k];
            FieldLoc L2;
            locate(c, f32, L2);
            std::vector<uint8_t> b2;
            TRY(copyArray(c, *L2.arr, b2, 16, true));
            float4* q = (float4*)b2.data();
            forEachParticle(c, [&](size_t h, size_t s) { ((float*)&q[s])[L2.comp] = hf[h]; });
            TRY(copyArray(c, *L2.arr, b2, 16, false));
        } else {
            forEachParticle(c, [&](size_t h, size_t s) { ((double*)hostBuf)[h] = d[s]; });
        }
        return ASTRAEA_OK;
    }
    void* dev = posPending ? (void*)c->dc.d.posNew : *L.arr;
    const size_t eb = elemBytes(L);
    std::vector<uint8_t> buf;
    TRY(copyArray(c, dev, buf, eb, true));
    uint8_t* raw = buf.data();
    if (L.kind == kF32) {
        forEachParticle(c, [&](size_t h, size_t s) {
            float* e = (float*)(raw + s * eb) + L.comp;
            if (upload) *e = ((const float*)hostBuf)[h]; else ((float*)hostBuf)[h] = *e;
        });
    } else if (L.kind == kU8 && L.comps == 0) {
        forEachParticle(c, [&](size_t h, size_t s) {
            if (upload) raw[s] = ((const uint8_t*)hostBuf)[h]; else ((uint8_t*)hostBuf)[h] = raw[s];
        });
    } else if (L.kind == kU8) {
        forEachParticle(c, [&](size_t h, size_t s) {
            uint32_t* m = (uint32_t*)(raw + s * 4);
            if (upload) {
                const uint32_t v = ((const uint8_t*)hostBuf)[h];
                *m = (*m & ~(0xffu << (8 * L.comp))) | (v << (8 * L.comp));
            } else {
                ((uint8_t*)hostBuf)[h] = (uint8_t)((*m >> (8 * L.comp)) & 0xffu);
            }
        });
    } else {  // kU32: original id, or a float count stored in a float4
        forEachParticle(c, [&](size_t h, size_t s) {
            if (L.comps == 1) {
                uint32_t* e = (uint32_t*)(raw + s * 4);
                if (upload) *e = ((const uint32_t*)hostBuf)[h]; else ((uint32_t*)hostBuf)[h] = *e;
            } else {
                const float v = ((float*)(raw + s * eb))[L.comp];
                ((uint32_t*)hostBuf)[h] = (uint32_t)v;
            }
        });
    }
    if (upload) {
        TRY(copyArray(c, dev, buf, eb, false));
        if (c->dc.fp64 && ((f >= ASTRAEA_F_POS_X && f <= ASTRAEA_F_POS_Z) || (f >= ASTRAEA_F_VEL_X && f <= ASTRAEA_F_VEL_Z))) {
            // the FP64 masters follow an FP32 upload
            const bool isPos = f <= ASTRAEA_F_POS_Z;
            const int comp = isPos ? f - ASTRAEA_F_POS_X : f - ASTRAEA_F_VEL_X;
            double* base = (isPos ? c->dc.cur.p64 : c->dc.cur.v64) + (size_t)comp * c->capSlots;
            std::vector<uint8_t> b64;
            TRY(copyArray(c, base, b64, sizeof(double), true));
            double* d = (double*)b64.data();
            forEachParticle(c, [&](size_t h, size_t s) { d[s] = (double)((const float*)hostBuf)[h]; });
            TRY(copyArray(c, base, b64, sizeof(double), false));
        }
        // positions, masses and smoothing lengths shape the tree, its moments and the neighbour
        // table; the other fields leave them valid
        const bool shapesTree = (f >= ASTRAEA_F_POS_X && f <= ASTRAEA_F_POS_Z) || f == ASTRAEA_F_MASS ||
                                f == ASTRAEA_F_H || f == ASTRAEA_F_H0 || (f >= ASTRAEA_F_POS64_X && f <= ASTRAEA_F_POS64_Z);
        if (shapesTree) {
            c->treeValid = c->listValid = false;
            c->posNewPending = false;
        }
        c->dtRawValid = false;
    }
    return ASTRAEA_OK;
}

// ---- timing helpers: start/end events on the context stream, no host wait ----
void collectTimes(astraea_ctx* c) {
    if (c->pending.empty()) return;
    (void)hipStreamSynchronize(c->stream);
    for (const auto& r : c->pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, c->evPool[r.a], c->evPool[r.b]) == hipSuccess) {
            c->stageMs[r.stage] += ms;
            c->stageN[r.stage] += 1;
        }
    }
    c->pending.clear();
    c->evNext = 0;
}
int takeEvent(astraea_ctx* c) {
    if (c->evNext >= (int)c->evPool.size()) collectTimes(c);
    return c->evNext++;
}
void stageMark(astraea_ctx* c, int stage) {
    if (!c->timing || c->evPool.empty()) return;
    const int e = takeEvent(c);
    (void)hipEventRecord(c->evPool[e], c->stream);
    c->openEv[stage] = e;
}
void stageClose(astraea_ctx* c, int stage) {
    if (!c->timing || c->evPool.empty()) return;
    const int e = takeEvent(c);
    (void)hipEventRecord(c->evPool[e], c->stream);
    c->pending.push_back({stage, c->openEv[stage], e});
}

void swapSets(astraea_ctx* c) {
    std::swap(c->dc.cur, c->dc.alt);
    c->parity ^= 1;
}

// Sort -> permute/Karras -> moments -> neighbour table.  posNew and the bounds must be ready.
void enqueueTree(astraea_ctx* c, bool timed, bool gravitySide = false) {
    hipStream_t s = c->stream;
    if (timed) stageMark(c, ASTRAEA_STAGE_MORTON);
    launchBoundsReduce(c->dc, s);
    launchMorton(sortInput(c->dc), s);
    if (timed) { stageClose(c, ASTRAEA_STAGE_MORTON); stageMark(c, ASTRAEA_STAGE_SORT); }
    sortPairs(c->dc, s);
    if (timed) { stageClose(c, ASTRAEA_STAGE_SORT); stageMark(c, ASTRAEA_STAGE_PERMUTE_TREE); }
    launchPermuteKarras(c->dc, s);
    swapSets(c);
    if (timed) { stageClose(c, ASTRAEA_STAGE_PERMUTE_TREE); stageMark(c, ASTRAEA_STAGE_MOMENTS); }
    launchMoments(c->dc, s);
    if (timed) stageClose(c, ASTRAEA_STAGE_MOMENTS);
    if (gravitySide) {
        // gravity needs the tree and the positions only: it runs on the side stream from here
        (void)hipEventRecord(c->fork, s);
        (void)hipStreamWaitEvent(c->side, c->fork, 0);
        launchGravity(c->dc, c->side);
        (void)hipEventRecord(c->join, c->side);
    }
    if (timed) stageMark(c, ASTRAEA_STAGE_NEIGHBOURS);
    launchNeighbours(c->dc, s);
    if (timed) stageClose(c, ASTRAEA_STAGE_NEIGHBOURS);
    c->treeValid = c->listValid = true;
    c->posNewPending = false;
}

// One full KDK step (Simulation::step, engine/simulation.cpp:1721-1732), fully asynchronous.
void enqueueStep(astraea_ctx* c, float maxDt, bool timed) {
    hipStream_t s = c->stream;
    if (timed) stageMark(c, ASTRAEA_STAGE_DT);
    launchStepBegin(c->dc, s, maxDt);
    if (timed) { stageClose(c, ASTRAEA_STAGE_DT); stageMark(c, ASTRAEA_STAGE_KICK1); }
    launchKick(c->dc, s, false, false);
    if (timed) { stageClose(c, ASTRAEA_STAGE_KICK1); stageMark(c, ASTRAEA_STAGE_DRIFT); }
    launchDrift(c->dc, s);
    if (timed) stageClose(c, ASTRAEA_STAGE_DRIFT);
    // untimed, gravity overlaps the neighbour build and pass 1 (the two fill each other's idle
    // CUs at the ends of their launches); timed, the stages run one after another
    enqueueTree(c, timed, !timed);
    if (timed) stageMark(c, ASTRAEA_STAGE_PASS1);
    launchPass1(c->dc, s);
    if (timed) { stageClose(c, ASTRAEA_STAGE_PASS1); stageMark(c, ASTRAEA_STAGE_GRAVITY); }
    if (timed) launchGravity(c->dc, s);
    else (void)hipStreamWaitEvent(s, c->join, 0);
    if (timed) { stageClose(c, ASTRAEA_STAGE_GRAVITY); stageMark(c, ASTRAEA_STAGE_PASS2); }
    launchPass2(c->dc, s);
    if (timed) { stageClose(c, ASTRAEA_STAGE_PASS2); stageMark(c, ASTRAEA_STAGE_KICK2); }
    launchKick(c->dc, s, true, false);
    launchStepEnd(c->dc, s);
    if (timed) stageClose(c, ASTRAEA_STAGE_KICK2);
    c->dtRawValid = true;
}

// Queues `steps` full steps, eagerly or (pairs of steps) through the captured graph, recording
// them for settle.
int queueSteps(astraea_ctx* c, float maxDt, uint32_t steps, bool useGraph) {
    uint32_t left = steps;
    if (useGraph && !c->timing) {
        if (c->graph && (c->graphMaxDt != maxDt || c->graphParity != c->parity)) freeGraph(c);
        if (left >= 2 && !c->graph) {
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            const int parity0 = c->parity;
            enqueueStep(c, maxDt, false);
            enqueueStep(c, maxDt, false);
            const hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (e != hipSuccess) return fail(ASTRAEA_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
            const hipError_t ei = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) return fail(ASTRAEA_E_HIP, "graph instantiate failed: %s", hipGetErrorString(ei));
            c->graphMaxDt = maxDt;
            c->graphParity = parity0;
            // the capture recorded two steps without running them; the two swaps restored cur/alt
            c->parity = parity0;
        }
        while (left >= 2 && c->graph) {
            HIP_TRY(hipGraphLaunch(c->graph, c->stream));
            c->stepsQueued += 2;
            c->unsettled.insert(c->unsettled.end(), 2, maxDt);
            left -= 2;
        }
    }
    while (left > 0) {
        enqueueStep(c, maxDt, c->timing);
        c->stepsQueued += 1;
        c->unsettled.push_back(maxDt);
        --left;
    }
    c->dtRawValid = true;
    return ASTRAEA_OK;
}

int requireLoaded(astraea_ctx* c) {
    if (!c) return fail(ASTRAEA_E_INVALID, "null context");
    if (!c->loaded) return fail(ASTRAEA_E_STATE, "no particle layout: call astraea_set_layout first");
    return ASTRAEA_OK;
}

int checkLaunch(astraea_ctx* c) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ASTRAEA_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
    (void)c;
    return ASTRAEA_OK;
}

// Waits for the tree just enqueued; when a particle has more neighbour candidates than the table
// holds, widens the table and rebuilds it from the same tree.  Used where the caller is not
// inside a run of steps (a load or an explicit tree build), so no step is ever lost to it.
int fitNeighbourTable(astraea_ctx* c) {
    for (int round = 0;; ++round) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        unsigned long long v[8], q[4];
        HIP_TRY(hipMemcpy(v, c->dc.cnt.v, sizeof(v), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(q, c->dc.nb.q, sizeof(q), hipMemcpyDeviceToHost));
        if (v[kCntHold] == 0ull) return ASTRAEA_OK;
        // clear the hold and the overflow flags and sizes (the counts 0, 1 and 4 stay)
        HIP_TRY(hipMemset(c->dc.cnt.v + kCntHold, 0, 2 * sizeof(unsigned long long)));
        HIP_TRY(hipMemset(c->dc.cnt.v + 5, 0, 3 * sizeof(unsigned long long)));
        if (v[6] != 0ull) {
            if (c->dc.tune.buildStack < 1024u) {
                // the shorter stack of a tuned variant filled: the full one from now on
                c->dc.tune.buildWaves = 4u;
                c->dc.tune.buildStack = 1024u;
                freeGraph(c);
                launchNeighbours(c->dc, c->stream);
                continue;
            }
            c->listValid = false;
            return fail(ASTRAEA_E_OVERFLOW, "neighbour traversal stack overflow");
        }
        if (round > 2 || v[5] > (1ull << 16)) {
            c->listValid = false;
            return fail(ASTRAEA_E_OVERFLOW,
                        "a wavefront's neighbours form a union of %llu ids, more than the 65536 a table can address",
                        (unsigned long long)v[5]);
        }
        // the spill pool was too small: q[1] holds what every listed wavefront reserved
        TRY(allocSpillPool(c, std::max<unsigned long long>(q[1] + q[1] / 4, 2 * c->dc.nb.poolBytes)));
        launchNeighbours(c->dc, c->stream);
    }
}

// Waits for the work queued on the context stream.  When the neighbour build of a queued step held
// the step sequence (its table came out incomplete: stepHeld), the state is that of the held step
// after its sort, with the step's dt: the table is widened and built again from the same tree, the
// rest of that step runs (pass 1, gravity, pass 2, second kick, epilogue), and then the steps that
// were queued after it, again until none is held.  The result is the run a table wide enough from
// the start would have given.
int settle(astraea_ctx* c) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (!c->loaded) return ASTRAEA_OK;
    for (;;) {
        unsigned long long v[8];
        HIP_TRY(hipMemcpy(v, c->dc.cnt.v, sizeof(v), hipMemcpyDeviceToHost));
        if (v[kCntHold] == 0ull) {
            c->unsettled.clear();
            return ASTRAEA_OK;
        }
        // the held step and the steps queued after it (all from astraea_step: every other path
        // fits its table before it goes on)
        const uint64_t lost = c->stepsQueued - v[kCntSteps];
        if (lost == 0ull || lost > c->unsettled.size()) {
            HIP_TRY(hipMemset(c->dc.cnt.v + kCntHold, 0, 2 * sizeof(unsigned long long)));
            HIP_TRY(hipMemset(c->dc.cnt.v + 5, 0, 3 * sizeof(unsigned long long)));
            c->stepsQueued = v[kCntSteps];
            c->unsettled.clear();
            c->listValid = false;
            return fail(ASTRAEA_E_OVERFLOW, "neighbour table overflow outside a step: evaluate the forces again");
        }
        std::vector<float> rest(c->unsettled.end() - (ptrdiff_t)lost, c->unsettled.end());
        c->unsettled.clear();
        c->stepsQueued -= lost;
        const int rc = fitNeighbourTable(c);
        if (rc != ASTRAEA_OK) return rc;  // listValid is false: the forces must be evaluated again
        launchPass1(c->dc, c->stream);
        launchGravity(c->dc, c->stream);
        launchPass2(c->dc, c->stream);
        launchKick(c->dc, c->stream, true, false);
        launchStepEnd(c->dc, c->stream);
        c->stepsQueued += 1;
        c->stepsRedone += lost;
        c->dtRawValid = true;
        TRY(checkLaunch(c));
        for (size_t k = 1; k < rest.size(); ++k) TRY(queueSteps(c, rest[k], 1, false));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
}

// settle, when steps may be pending; every entry point but astraea_step starts with it.
int settleQueued(astraea_ctx* c) { return c->unsettled.empty() ? ASTRAEA_OK : settle(c); }

}  // namespace

extern "C" {

size_t astraea_field_size(int f) {
    if (f < 0 || f >= ASTRAEA_F_COUNT_) return 0;
    if (f == ASTRAEA_F_MATERIAL || f == ASTRAEA_F_BODY || f == ASTRAEA_F_TRACER || f == ASTRAEA_F_PHASE) return 1;
    if (f >= ASTRAEA_F_POS64_X && f <= ASTRAEA_F_VEL64_Z) return 8;
    return 4;
}

const char* astraea_last_error(void) { return gLastError.c_str(); }
int astraea_abi_version(void) { return ASTRAEA_ABI_VERSION; }

void astraea_default_params(astraea_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->theta = 0.5f;
    p->soft_factor = 0.8f;
    p->cfl_factor = 0.25f;
    p->gamma_hhe = 1.4f;
    p->cooling_min_temp_k = 0.0f;
    p->strength_mask = 0x17u;
    p->precision = ASTRAEA_PREC_FP32;
    p->group_size = 32u;  // the reference's warp groups
}

int astraea_create(astraea_ctx** out, int device, uint32_t capacity, uint32_t maxSegments) {
    if (!out) return fail(ASTRAEA_E_INVALID, "null output pointer");
    *out = nullptr;
    if (capacity == 0 || capacity > ASTRAEA_MAX_CAPACITY)
        return fail(ASTRAEA_E_INVALID, "capacity must be in [1, %u]", (unsigned)ASTRAEA_MAX_CAPACITY);
    if (maxSegments == 0 || maxSegments > 256) return fail(ASTRAEA_E_INVALID, "max_segments must be in [1, 256]");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(ASTRAEA_E_INVALID, "device %d out of range (%d devices)", device, ndev);
    HIP_TRY(hipSetDevice(device));
    auto* c = new astraea_ctx();
    c->device = device;
    c->capSlots = roundUp(capacity, ASTRAEA_SEGMENT_ALIGN) + ASTRAEA_SEGMENT_ALIGN * maxSegments;
    c->maxSeg = maxSegments;
    astraea_default_params(&c->params);
    syncDerivedFlags(c);
    auto cleanup = [&](int r) {
        astraea_destroy(c);
        return r;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(ASTRAEA_E_HIP, "hipStreamCreate failed"));
    if (hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->join, hipEventDisableTiming) != hipSuccess)
        return cleanup(fail(ASTRAEA_E_HIP, "hipEventCreate failed"));
    {
        const int rs = scratchAddress(device, &c->scratchDev);
        if (rs != ASTRAEA_OK) return cleanup(rs);
    }
    const uint32_t C = c->capSlots;
    DeviceCtx& d = c->dc;
    int r = ASTRAEA_OK;
    auto A = [&](int x) { if (r == ASTRAEA_OK) r = x; };
    A(allocPersist(c, d.cur, C));
    A(allocPersist(c, d.alt, C));
    A(devAlloc(c, d.d.posNew, C));
    A(devAlloc(c, d.d.grav, C));
    A(devAlloc(c, d.d.rate, C));
    A(devAlloc(c, d.d.rate2, C));
    A(devAlloc(c, d.d.dS0, C));
    A(devAlloc(c, d.d.dS1, C));
    A(devAlloc(c, d.d.eos, C));
    A(devAlloc(c, d.d.phase, C));
    A(devAlloc(c, d.d.p1a, C));
    A(devAlloc(c, d.d.p1b, C));
    A(devAlloc(c, d.d.nb, C));
    A(devAlloc(c, d.d.sr0, C));
    A(devAlloc(c, d.d.sr1, C));
    A(devAlloc(c, d.t.child, C));
    A(devAlloc(c, d.t.parentI, C));
    A(devAlloc(c, d.t.parentL, C));
    A(devAlloc(c, d.t.visited, C));
    A(devAlloc(c, d.t.other, C));
    A(devAlloc(c, d.t.cm, C));
    A(devAlloc(c, d.t.bmin, C));
    A(devAlloc(c, d.t.bmax, C));
    A(devAlloc(c, d.nb.count, C));
    A(devAlloc(c, d.nb.waveSum, C / kWave));
    A(devAlloc(c, d.nb.waveMax, C / kWave));
    A(devAlloc(c, d.nb.waveNeed, C / kWave));
    A(devAlloc(c, d.nb.unionLen, C / kWave));
    A(devAlloc(c, d.nb.tab, C / kWave));
    A(devAlloc(c, d.nb.spill, C / kWave));
    A(devAlloc(c, d.nb.q, 4));
    A(devAlloc(c, d.seg.off, maxSegments + 1));
    A(devAlloc(c, d.seg.n, maxSegments));
    A(devAlloc(c, d.seg.slotSeg, C));
    A(devAlloc(c, d.ss.bounds, 6 * (size_t)maxSegments));
    A(devAlloc(c, d.d.blockBounds, 6 * (size_t)(C / kBlock)));
    A(devAlloc(c, d.ss.dtRawBits, maxSegments));
    A(devAlloc(c, d.ss.dt, maxSegments));
    A(devAlloc(c, d.ss.dtPrev, maxSegments));
    A(devAlloc(c, d.ss.radSum, maxSegments));
    A(devAlloc(c, d.ss.radPrev, maxSegments));
    A(devAlloc(c, d.ss.erad, maxSegments));
    A(devAlloc(c, d.ss.simTime, maxSegments));
    A(devAlloc(c, d.ss.living, maxSegments));
    A(devAlloc(c, d.cnt.v, 8));
    // tiles are cut along segment boundaries: at most one short tile per segment beyond C / tile
    const uint32_t tiles = C / kSortTile + maxSegments + 1u;
    A(devAlloc(c, d.sb.keyA, C));
    A(devAlloc(c, d.sb.keyB, C));
    A(devAlloc(c, d.sb.valA, C));
    A(devAlloc(c, d.sb.valB, C));
    A(devAlloc(c, d.sb.counts, kSortMaxDigits * (size_t)tiles));
    A(devAlloc(c, d.sb.tilePrefix, kSortMaxDigits * (size_t)tiles));
    A(devAlloc(c, d.sb.digitTotal, kSortMaxDigits * (size_t)maxSegments));
    A(devAlloc(c, d.sb.tileStart, (size_t)tiles + 1u));
    A(devAlloc(c, d.sb.tileSeg, tiles));
    A(devAlloc(c, d.sb.segTile0, (size_t)maxSegments + 1u));
    d.sb.tilesCap = tiles;
    d.sb.numTiles = 0;
    A(devAlloc(c, c->dEnergy, (size_t)ASTRAEA_ENERGY_STRIDE * maxSegments));
    A(devAlloc(c, c->dClass, (size_t)ASTRAEA_CLASS_STRIDE * maxSegments));
    A(devAlloc(c, c->dRem, (size_t)24 * maxSegments));
    A(devAlloc(c, c->dSeed, maxSegments));
    if (r == ASTRAEA_OK) r = allocNeighbourList(c, ASTRAEA_DEFAULT_NEIGHBOUR_STRIDE);
    if (r == ASTRAEA_OK) r = allocSpillPool(c, 1ull << 20);
    if (r != ASTRAEA_OK) return cleanup(r);
    d.stride64 = C;
    if (hipMemsetAsync(d.cnt.v, 0, 8 * sizeof(unsigned long long), c->stream) != hipSuccess)
        return cleanup(fail(ASTRAEA_E_HIP, "hipMemset failed"));
    if (hipStreamSynchronize(c->stream) != hipSuccess) return cleanup(fail(ASTRAEA_E_HIP, "stream sync failed"));
    *out = c;
    return ASTRAEA_OK;
}

int astraea_destroy(astraea_ctx* c) {
    if (!c) return ASTRAEA_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    freeGraph(c);
    for (void* p : c->allocs) (void)hipFree(p);
    for (hipEvent_t e : c->evPool) (void)hipEventDestroy(e);
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->join) (void)hipEventDestroy(c->join);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return ASTRAEA_OK;
}

void* astraea_stream(astraea_ctx* c) { return c ? (void*)c->stream : nullptr; }

int astraea_set_params(astraea_ctx* c, const astraea_params* p) {
    if (!c || !p) return fail(ASTRAEA_E_INVALID, "null argument");
    if (!(p->theta >= 0.0f) || !(p->soft_factor >= 0.0f) || !(p->cfl_factor > 0.0f))
        return fail(ASTRAEA_E_INVALID, "theta, soft_factor must be >= 0 and cfl_factor > 0");
    if (p->precision != ASTRAEA_PREC_FP32 && p->precision != ASTRAEA_PREC_FP64_ACCUM)
        return fail(ASTRAEA_E_INVALID, "unknown precision mode %u", p->precision);
    if (p->group_size != 0u && p->group_size != 32u && p->group_size != 64u)
        return fail(ASTRAEA_E_INVALID, "group_size %u is not 0, 32 or 64", p->group_size);
    if (c->loaded) TRY(settleQueued(c));
    const bool wasFp64 = c->params.precision == ASTRAEA_PREC_FP64_ACCUM;
    c->params = *p;
    syncDerivedFlags(c);
    freeGraph(c);
    if (c->loaded && !wasFp64 && c->dc.fp64) {
        // seed the FP64 masters from the FP32 state
        for (int comp = 0; comp < 3; ++comp) {
            const int fp = ASTRAEA_F_POS_X + comp, fv = ASTRAEA_F_VEL_X + comp;
            size_t n = 0;
            for (uint32_t s = 0; s < c->numSeg; ++s) n += c->counts[s];
            std::vector<float> tmp(n);
            std::vector<double> t64(n);
            TRY(transferField(c, fp, tmp.data(), false));
            for (size_t k = 0; k < n; ++k) t64[k] = tmp[k];
            TRY(transferField(c, ASTRAEA_F_POS64_X + comp, t64.data(), true));
            TRY(transferField(c, fv, tmp.data(), false));
            for (size_t k = 0; k < n; ++k) t64[k] = tmp[k];
            TRY(transferField(c, ASTRAEA_F_VEL64_X + comp, t64.data(), true));
        }
    }
    return ASTRAEA_OK;
}

int astraea_get_params(astraea_ctx* c, astraea_params* out) {
    if (!c || !out) return fail(ASTRAEA_E_INVALID, "null argument");
    *out = c->params;
    return ASTRAEA_OK;
}

int astraea_set_layout(astraea_ctx* c, uint32_t ns, const uint32_t* counts) {
    if (!c) return fail(ASTRAEA_E_INVALID, "null context");
    if (ns == 0 || ns > c->maxSeg || !counts) return fail(ASTRAEA_E_INVALID, "num_segments must be in [1, %u]", c->maxSeg);
    uint64_t total = 0, slots = 0;
    std::vector<uint32_t> offs(ns + 1, 0);
    for (uint32_t s = 0; s < ns; ++s) {
        offs[s] = (uint32_t)slots;
        total += counts[s];
        slots += roundUp(counts[s], ASTRAEA_SEGMENT_ALIGN);
    }
    offs[ns] = (uint32_t)slots;
    if (total == 0) return fail(ASTRAEA_E_INVALID, "the layout holds no particles (reference: totalParticles == 0)");
    if (slots > c->capSlots)
        return fail(ASTRAEA_E_CAPACITY, "layout needs %llu slots, capacity is %u", (unsigned long long)slots, c->capSlots);
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    freeGraph(c);
    c->numSeg = ns;
    c->counts.assign(counts, counts + ns);
    c->tracerCount = 0;  // the layout clears the tracer flags with the rest of the state
    c->offs = offs;
    c->slots = (uint32_t)std::max<uint64_t>(slots, ASTRAEA_SEGMENT_ALIGN);
    DeviceCtx& d = c->dc;
    d.slots = c->slots;
    d.blocks = c->slots / kBlock;
    d.seg.count = ns;
    std::vector<uint32_t> slotSeg(c->slots, 0);
    for (uint32_t s = 0; s < ns; ++s)
        for (uint32_t k = offs[s]; k < offs[s + 1]; ++k) slotSeg[k] = s;
    HIP_TRY(hipMemcpy(d.seg.off, offs.data(), sizeof(uint32_t) * (ns + 1), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.seg.n, counts, sizeof(uint32_t) * ns, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.seg.slotSeg, slotSeg.data(), sizeof(uint32_t) * c->slots, hipMemcpyHostToDevice));
    // the sort's tiles: each segment's slot range cut into tiles of kSortTile (the last may be short)
    {
        std::vector<uint32_t> tileStart, tileSeg, segTile0(ns + 1, 0);
        for (uint32_t s = 0; s < ns; ++s) {
            segTile0[s] = (uint32_t)tileSeg.size();
            for (uint32_t p = offs[s]; p < offs[s + 1]; p += kSortTile) {
                tileStart.push_back(p);
                tileSeg.push_back(s);
            }
        }
        segTile0[ns] = (uint32_t)tileSeg.size();
        tileStart.push_back(offs[ns]);
        if (tileSeg.size() > d.sb.tilesCap)
            return fail(ASTRAEA_E_CAPACITY, "%zu sort tiles, capacity %u", tileSeg.size(), d.sb.tilesCap);
        HIP_TRY(hipMemcpy(d.sb.tileStart, tileStart.data(), sizeof(uint32_t) * tileStart.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.sb.tileSeg, tileSeg.data(), sizeof(uint32_t) * tileSeg.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(d.sb.segTile0, segTile0.data(), sizeof(uint32_t) * segTile0.size(), hipMemcpyHostToDevice));
        d.sb.numTiles = (uint32_t)tileSeg.size();
    }
    // state as uploadHostStateToDevice leaves it (simulation.cpp:469-520): zeros, eps_fail = 1,
    // original id = index in the concatenation
    for (Persist* p : {&d.cur, &d.alt}) {
        HIP_TRY(hipMemset(p->pos, 0, sizeof(float4) * c->slots));
        HIP_TRY(hipMemset(p->vel, 0, sizeof(float4) * c->slots));
        HIP_TRY(hipMemset(p->hu, 0, sizeof(float4) * c->slots));
        HIP_TRY(hipMemset(p->str0, 0, sizeof(float4) * c->slots));
        HIP_TRY(hipMemset(p->meta, 0, sizeof(uint32_t) * c->slots));
        HIP_TRY(hipMemset(p->p64, 0, sizeof(double) * 3 * (size_t)c->capSlots));
        HIP_TRY(hipMemset(p->v64, 0, sizeof(double) * 3 * (size_t)c->capSlots));
    }
    std::vector<float4> s1(c->slots, make_float4(0.0f, 0.0f, 0.0f, 1.0f));
    HIP_TRY(hipMemcpy(d.cur.str1, s1.data(), sizeof(float4) * c->slots, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d.alt.str1, s1.data(), sizeof(float4) * c->slots, hipMemcpyHostToDevice));
    std::vector<uint32_t> oid(c->slots, 0xffffffffu);
    {
        uint32_t h = 0;
        for (uint32_t s = 0; s < ns; ++s)
            for (uint32_t k = 0; k < counts[s]; ++k) oid[offs[s] + k] = h++;
    }
    HIP_TRY(hipMemcpy(d.cur.oid, oid.data(), sizeof(uint32_t) * c->slots, hipMemcpyHostToDevice));
    for (float4* a : {d.d.posNew, d.d.grav, d.d.rate, d.d.rate2, d.d.dS0, d.d.eos, d.d.p1a, d.d.p1b, d.d.nb, d.d.sr0})
        HIP_TRY(hipMemset(a, 0, sizeof(float4) * c->slots));
    HIP_TRY(hipMemset(d.d.dS1, 0, sizeof(float2) * c->slots));
    HIP_TRY(hipMemset(d.d.sr1, 0, sizeof(float2) * c->slots));
    HIP_TRY(hipMemset(d.d.phase, 0, c->slots));
    HIP_TRY(hipMemset(d.nb.count, 0, sizeof(uint2) * c->slots));
    HIP_TRY(hipMemset(d.nb.waveSum, 0, sizeof(uint32_t) * (c->slots / kWave)));
    HIP_TRY(hipMemset(d.nb.waveMax, 0, sizeof(uint32_t) * (c->slots / kWave)));
    HIP_TRY(hipMemset(d.nb.waveNeed, 0, sizeof(uint32_t) * (c->slots / kWave)));
    HIP_TRY(hipMemset(d.nb.q, 0, 4 * sizeof(unsigned long long)));
    HIP_TRY(hipMemset(d.nb.unionLen, 0, sizeof(uint32_t) * (c->slots / kWave)));
    HIP_TRY(hipMemset(d.ss.dtPrev, 0, sizeof(float) * ns));
    HIP_TRY(hipMemset(d.ss.dt, 0, sizeof(float) * ns));
    HIP_TRY(hipMemset(d.ss.erad, 0, sizeof(double) * ns));
    HIP_TRY(hipMemset(d.ss.simTime, 0, sizeof(double) * ns));
    HIP_TRY(hipMemset(d.ss.radPrev, 0, sizeof(double) * ns));
    HIP_TRY(hipMemset(d.ss.radSum, 0, sizeof(double) * ns));
    HIP_TRY(hipMemset(d.ss.living, 0, sizeof(uint32_t) * ns));
    HIP_TRY(hipMemset(d.ss.dtRawBits, 0x7f, sizeof(uint32_t) * ns));
    HIP_TRY(hipMemset(d.cnt.v, 0, 8 * sizeof(unsigned long long)));
    HIP_TRY(hipDeviceSynchronize());
    c->stepsQueued = 0;
    c->unsettled.clear();
    c->loaded = true;
    c->treeValid = c->listValid = c->dtRawValid = c->posNewPending = false;
    return ASTRAEA_OK;
}

int astraea_num_particles(astraea_ctx* c, uint32_t* out) {
    if (!c || !out) return fail(ASTRAEA_E_INVALID, "null argument");
    uint32_t n = 0;
    for (uint32_t v : c->counts) n += v;
    *out = n;
    return ASTRAEA_OK;
}

int astraea_upload_field(astraea_ctx* c, int f, const void* src) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    HIP_TRY(hipSetDevice(c->device));
    freeGraph(c);
    return transferField(c, f, const_cast<void*>(src), true);
}

int astraea_download_field(astraea_ctx* c, int f, void* dst) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    HIP_TRY(hipSetDevice(c->device));
    return transferField(c, f, dst, false);
}

int astraea_clear_stats(astraea_ctx* c) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    launchClear(c->dc, c->stream);
    c->dtRawValid = false;
    return checkLaunch(c);
}


int astraea_sort_and_build_bvh(astraea_ctx* c) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    if (!c->posNewPending) {
        launchResetBounds(c->dc, c->stream);
        launchPrepareSort(c->dc, c->stream);
    }
    enqueueTree(c, false);
    TRY(checkLaunch(c));
    TRY(fitNeighbourTable(c));
    c->dtRawValid = false;
    return checkLaunch(c);
}

int astraea_sph_pass1(astraea_ctx* c) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    if (!c->listValid) return fail(ASTRAEA_E_STATE, "pass 1 needs a tree: call astraea_sort_and_build_bvh");
    launchPass1(c->dc, c->stream);
    c->dtRawValid = false;
    return checkLaunch(c);
}

int astraea_evaluate_eos(astraea_ctx* c) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    launchEos(c->dc, c->stream);
    c->dtRawValid = false;
    return checkLaunch(c);
}

int astraea_gravity(astraea_ctx* c) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    if (!c->treeValid) return fail(ASTRAEA_E_STATE, "gravity needs a tree: call astraea_sort_and_build_bvh");
    launchGravity(c->dc, c->stream);
    c->dtRawValid = false;
    return checkLaunch(c);
}

int astraea_sph_pass2(astraea_ctx* c) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    if (!c->listValid) return fail(ASTRAEA_E_STATE, "pass 2 needs a tree: call astraea_sort_and_build_bvh");
    launchPass2(c->dc, c->stream);
    c->dtRawValid = false;
    return checkLaunch(c);
}

int astraea_reduce_dt(astraea_ctx* c, float maxDt, float* out) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    if (!(maxDt > 0.0f)) return fail(ASTRAEA_E_INVALID, "max_dt must be > 0");
    launchDtCandidates(c->dc, c->stream);
    launchStepBeginNoBounds(c->dc, c->stream, maxDt);
    TRY(checkLaunch(c));
    if (out) {
        HIP_TRY(hipMemcpyAsync(out, c->dc.ss.dt, sizeof(float) * c->numSeg, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return ASTRAEA_OK;
}

int astraea_kick_half(astraea_ctx* c, int clamp) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    launchKick(c->dc, c->stream, false, clamp != 0);
    c->dtRawValid = false;
    return checkLaunch(c);
}

int astraea_drift(astraea_ctx* c) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    if (c->posNewPending) return fail(ASTRAEA_E_STATE, "a drift must be followed by astraea_sort_and_build_bvh");
    launchResetBounds(c->dc, c->stream);
    launchDrift(c->dc, c->stream);
    c->posNewPending = true;
    c->dtRawValid = false;
    return checkLaunch(c);
}

int astraea_evaluate_initial_forces(astraea_ctx* c) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    HIP_TRY(hipSetDevice(c->device));
    hipStream_t s = c->stream;
    launchClear(c->dc, s);
    launchResetBounds(c->dc, s);
    launchPrepareSort(c->dc, s);
    enqueueTree(c, false);
    TRY(checkLaunch(c));
    TRY(fitNeighbourTable(c));
    launchPass1(c->dc, s);
    launchGravity(c->dc, s);
    launchPass2(c->dc, s);
    launchDtCandidates(c->dc, s);
    launchInitIntegrator(c->dc, s);
    c->dtRawValid = true;
    return checkLaunch(c);
}

int astraea_step(astraea_ctx* c, float maxDt, uint32_t steps, int useGraph) {
    TRY(requireLoaded(c));
    if (!(maxDt > 0.0f)) return fail(ASTRAEA_E_INVALID, "max_dt must be > 0");
    if (!c->listValid || c->posNewPending)
        return fail(ASTRAEA_E_STATE, "step needs evaluated forces: call astraea_evaluate_initial_forces");
    HIP_TRY(hipSetDevice(c->device));
    if (!c->dtRawValid) {
        launchDtCandidates(c->dc, c->stream);
        c->dtRawValid = true;
    }
    TRY(queueSteps(c, maxDt, steps, useGraph != 0));
    return checkLaunch(c);
}

int astraea_synchronize(astraea_ctx* c) {
    if (!c) return fail(ASTRAEA_E_INVALID, "null context");
    HIP_TRY(hipSetDevice(c->device));
    return settle(c);
}

int astraea_get_dt(astraea_ctx* c, float* out) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    if (!out) return fail(ASTRAEA_E_INVALID, "null output");
    HIP_TRY(hipMemcpyAsync(out, c->dc.ss.dt, sizeof(float) * c->numSeg, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ASTRAEA_OK;
}

int astraea_energy(astraea_ctx* c, double* out) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    if (!out) return fail(ASTRAEA_E_INVALID, "null output");
    HIP_TRY(hipSetDevice(c->device));
    launchEnergy(c->dc, c->stream, c->dEnergy);
    TRY(checkLaunch(c));
    const uint32_t S = c->numSeg;
    std::vector<double> erad(S), t(S);
    HIP_TRY(hipMemcpyAsync(out, c->dEnergy, sizeof(double) * ASTRAEA_ENERGY_STRIDE * S, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(erad.data(), c->dc.ss.erad, sizeof(double) * S, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(t.data(), c->dc.ss.simTime, sizeof(double) * S, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (uint32_t s = 0; s < S; ++s) {
        double* o = out + (size_t)s * ASTRAEA_ENERGY_STRIDE;
        const double m = o[ASTRAEA_EN_MASS];
        if (m > 0.0)
            for (int k = ASTRAEA_EN_COM_X; k <= ASTRAEA_EN_COM_Z; ++k) o[k] /= m;
        o[ASTRAEA_EN_RADIATED] = erad[s];
        o[ASTRAEA_EN_SIM_TIME] = t[s];
    }
    return ASTRAEA_OK;
}

int astraea_classify(astraea_ctx* c, const double* targetRadius, double* out) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    if (!targetRadius || !out) return fail(ASTRAEA_E_INVALID, "null argument");
    const uint32_t S = c->numSeg;
    for (uint32_t s = 0; s < S; ++s)
        if (!(targetRadius[s] >= 0.0) || !std::isfinite(targetRadius[s]))
            return fail(ASTRAEA_E_INVALID, "target radius of segment %u is %g", s, targetRadius[s]);
    HIP_TRY(hipSetDevice(c->device));
    constexpr int W = ASTRAEA_CLASS_STRIDE;
    double* dGuess = c->dRem;
    double* dAcc = c->dRem + 8 * (size_t)S;
    double* dBody = c->dRem + 16 * (size_t)S;
    launchClassBodies(c->dc, c->stream, c->dClass, c->dSeed);
    TRY(checkLaunch(c));
    std::vector<double> first((size_t)W * S);
    HIP_TRY(hipMemcpyAsync(first.data(), c->dClass, sizeof(double) * W * S, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));

    // Remnant guess (simulation.cpp:1115-1123): the seed's state, mass max(m0, M/2), radius
    // max(R_target, 0.25).  rem[0] < 0 marks a segment that stopped iterating.
    std::vector<double> rem(8 * (size_t)S), guess(8 * (size_t)S), acc(8 * (size_t)S), body(8 * (size_t)S, 0.0);
    std::vector<int> iters(S, 0);
    std::vector<bool> active(S);
    for (uint32_t s = 0; s < S; ++s) {
        double* o = first.data() + (size_t)s * W;
        const double m0 = o[ASTRAEA_CL_BODY0_MASS], m1 = o[ASTRAEA_CL_BODY1_MASS];
        for (int b = 0; b < 2; ++b) {
            const double m = b ? m1 : m0;
            double* v = o + (b ? ASTRAEA_CL_BODY1_COM : ASTRAEA_CL_BODY0_COM);
            for (int k = 0; k < 6; ++k) v[k] = m > 0.0 ? v[k] / m : 0.0;
        }
        double* r = rem.data() + 8 * (size_t)s;
        r[0] = std::max(m0, 0.5 * (m0 + m1));
        for (int k = 0; k < 6; ++k) r[1 + k] = o[ASTRAEA_CL_SEED_X + k];
        r[7] = std::max(targetRadius[s], 0.25);
        active[s] = c->counts[s] > 0;
        double* b0 = body.data() + 8 * (size_t)s;
        for (int k = 0; k < 6; ++k) b0[k] = o[ASTRAEA_CL_BODY0_COM + k];
        b0[6] = targetRadius[s];
    }
    // Five iterations (simulation.cpp:1125-1178), all segments per launch.
    for (int it = 0; it < 5; ++it) {
        bool any = false;
        for (uint32_t s = 0; s < S; ++s) {
            for (int k = 0; k < 8; ++k) guess[8 * (size_t)s + k] = rem[8 * (size_t)s + k];
            if (!active[s]) guess[8 * (size_t)s] = -1.0;
            any = any || active[s];
        }
        if (!any) break;
        HIP_TRY(hipMemcpyAsync(dGuess, guess.data(), sizeof(double) * 8 * S, hipMemcpyHostToDevice, c->stream));
        launchRemnant(c->dc, c->stream, dGuess, dAcc);
        TRY(checkLaunch(c));
        HIP_TRY(hipMemcpyAsync(acc.data(), dAcc, sizeof(double) * 8 * S, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        for (uint32_t s = 0; s < S; ++s) {
            if (!active[s]) continue;
            const double* a = acc.data() + 8 * (size_t)s;
            double* r = rem.data() + 8 * (size_t)s;
            if (!(a[0] > 0.0)) {
                active[s] = false;
                continue;
            }
            r[0] = a[0];
            for (int k = 1; k < 7; ++k) r[k] = a[k] / a[0];
            r[7] = std::max(0.5 * targetRadius[s], std::sqrt(5.0 / 3.0) * std::sqrt(a[7] / a[0]));
            iters[s] += 1;
        }
    }
    // Final split (simulation.cpp:1180-1227) and the spin of body 0 (1094-1112).
    HIP_TRY(hipMemcpyAsync(dGuess, rem.data(), sizeof(double) * 8 * S, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dBody, body.data(), sizeof(double) * 8 * S, hipMemcpyHostToDevice, c->stream));
    launchClassify(c->dc, c->stream, dGuess, dBody, c->dClass);
    TRY(checkLaunch(c));
    HIP_TRY(hipMemcpyAsync(out, c->dClass, sizeof(double) * W * S, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (uint32_t s = 0; s < S; ++s) {
        double* o = out + (size_t)s * W;
        const double* f = first.data() + (size_t)s * W;
        for (int k = 0; k < ASTRAEA_CL_REM_MASS; ++k) o[k] = f[k];
        for (int k = 0; k < 8; ++k) o[ASTRAEA_CL_REM_MASS + k] = rem[8 * (size_t)s + k];
        o[ASTRAEA_CL_ITERATIONS] = iters[s];
    }
    return ASTRAEA_OK;
}

// One relaxation step (simulation.cpp:603-615): the KDK sequence with dt_max 0.02 PSU and the
// thermal clamp in both kicks.  The tree is fitted (the table widened if needed) before its use.
static int enqueueRelaxKdk(astraea_ctx* c) {
    hipStream_t s = c->stream;
    launchStepBegin(c->dc, s, 0.02f);
    launchKick(c->dc, s, false, true);
    launchDrift(c->dc, s);
    enqueueTree(c, false);
    TRY(checkLaunch(c));
    TRY(fitNeighbourTable(c));
    launchPass1(c->dc, s);
    launchGravity(c->dc, s);
    launchPass2(c->dc, s);
    launchKick(c->dc, s, true, true);
    launchStepEnd(c->dc, s);
    c->stepsQueued += 1;
    return checkLaunch(c);
}

int astraea_relax(astraea_ctx* c, uint32_t steps, const float* anchorTempK, const double* nominalRadius,
                  double* out) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    if (!nominalRadius || !out) return fail(ASTRAEA_E_INVALID, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    const uint32_t S = c->numSeg;
    uint32_t total = 0;
    for (uint32_t k = 0; k < S; ++k) total += c->counts[k];
    if (!c->dAnchor) TRY(devAlloc(c, c->dAnchor, c->capSlots));
    hipStream_t s = c->stream;
    double* sums = c->dRem;
    double* stats = c->dRem + 8 * (size_t)S;
    TRY(astraea_evaluate_initial_forces(c));
    if (anchorTempK)
        HIP_TRY(hipMemcpyAsync(c->dAnchor, anchorTempK, sizeof(float) * total, hipMemcpyHostToDevice, s));
    else
        launchAnchorFromEos(c->dc, s, c->dAnchor, total);
    for (uint32_t k = 0; k < steps; ++k) {
        TRY(enqueueRelaxKdk(c));
        // centre, damp, re-anchor, clear stress and damage (simulation.cpp:637-700)
        launchRelaxCorrect(c->dc, s, sums, c->dAnchor, total, 0.85);
        if (k + 1 < steps) {  // forces of the corrected state (simulation.cpp:702-706)
            launchEos(c->dc, s);
            launchGravity(c->dc, s);
            launchPass2(c->dc, s);
            launchDtCandidates(c->dc, s);
        }
        TRY(checkLaunch(c));
    }
    if (steps > 0) {  // four undamped steps (simulation.cpp:709-726)
        launchResetBounds(c->dc, s);
        launchPrepareSort(c->dc, s);
        enqueueTree(c, false);
        TRY(checkLaunch(c));
        TRY(fitNeighbourTable(c));
        launchPass1(c->dc, s);
        launchGravity(c->dc, s);
        launchPass2(c->dc, s);
        launchDtCandidates(c->dc, s);
        for (int k = 0; k < 4; ++k) TRY(enqueueRelaxKdk(c));
        launchEos(c->dc, s);
    }
    launchRelaxStats(c->dc, s, stats);
    TRY(checkLaunch(c));
    std::vector<double> st(8 * (size_t)S);
    HIP_TRY(hipMemcpyAsync(st.data(), stats, sizeof(double) * 8 * S, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (uint32_t k = 0; k < S; ++k) {  // simulation.cpp:729-754
        const double* q = st.data() + 8 * (size_t)k;
        const double vRms = q[0] > 0.0 ? std::sqrt(q[1] / q[0]) : 0.0;
        const double rScale = std::max(std::max(nominalRadius[k], std::sqrt(q[2])), 1.0e-3);
        const double vEsc = std::sqrt(2.0 * 66.7430 * q[0] / rScale);
        out[k] = vEsc > 0.0 ? vRms / vEsc : 0.0;
    }
    if (steps > 0) {  // re-anchor and clear once more (simulation.cpp:757-780)
        launchReanchor(c->dc, s, c->dAnchor, total);
        launchEos(c->dc, s);
        TRY(checkLaunch(c));
    }
    // the forces are those of the state before the last corrections: evaluate them again to step
    c->listValid = false;
    c->dtRawValid = false;
    HIP_TRY(hipStreamSynchronize(s));
    return ASTRAEA_OK;
}

static int ensureRenderBuffers(astraea_ctx* c) {
    if (c->dRender) return ASTRAEA_OK;
    TRY(devAlloc(c, c->dRender, 5 * (size_t)c->capSlots));
    TRY(devAlloc(c, c->dSpeed, c->capSlots));
    TRY(devAlloc(c, c->dCom, 6 * (size_t)c->maxSeg));
    TRY(devAlloc(c, c->dOutBase, c->maxSeg));
    TRY(devAlloc(c, c->dTracerOf, c->capSlots));
    return ASTRAEA_OK;
}

int astraea_pack_render(astraea_ctx* c, const double* comPsu, uint8_t* records, uint8_t* speed) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    if (!comPsu || !records) return fail(ASTRAEA_E_INVALID, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    TRY(ensureRenderBuffers(c));
    const uint32_t S = c->numSeg;
    std::vector<float> com(3 * (size_t)S);
    std::vector<uint32_t> base(S);
    uint32_t total = 0;
    for (uint32_t k = 0; k < S; ++k) {
        for (int a = 0; a < 3; ++a) com[3 * k + a] = (float)comPsu[3 * k + a];
        base[k] = total;
        total += c->counts[k];
    }
    HIP_TRY(hipMemcpyAsync(c->dCom, com.data(), sizeof(float) * 3 * S, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->dOutBase, base.data(), sizeof(uint32_t) * S, hipMemcpyHostToDevice, c->stream));
    launchPackRender(c->dc, c->stream, c->dCom, c->dOutBase, c->dRender, speed ? c->dSpeed : nullptr);
    TRY(checkLaunch(c));
    HIP_TRY(hipMemcpyAsync(records, c->dRender, 20 * (size_t)total, hipMemcpyDeviceToHost, c->stream));
    if (speed) HIP_TRY(hipMemcpyAsync(speed, c->dSpeed, total, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ASTRAEA_OK;
}

int astraea_pack_render_noalloc(astraea_ctx* c, const double* comPsu, uint8_t* recordsDev, uint8_t* speedDev) {
    TRY(requireLoaded(c));
    if (!comPsu || !recordsDev) return fail(ASTRAEA_E_INVALID, "null argument");
    const uint32_t S = c->numSeg;
    if (S > ASTRAEA_NOALLOC_MAX_SEGMENTS)
        return fail(ASTRAEA_E_INVALID, "%u segments: the allocation-free render takes at most %u", S,
                    (unsigned)ASTRAEA_NOALLOC_MAX_SEGMENTS);
    if (!c->scratchDev) return fail(ASTRAEA_E_STATE, "no scratch table on device %d", c->device);
    HIP_TRY(hipSetDevice(c->device));
    if (!c->unsettled.empty()) {
        // queued steps: finishing a held one would widen (allocate) the neighbour table
        HIP_TRY(hipStreamSynchronize(c->stream));
        unsigned long long hold = 0ull;
        HIP_TRY(hipMemcpy(&hold, c->dc.cnt.v + kCntHold, sizeof(hold), hipMemcpyDeviceToHost));
        if (hold != 0ull)
            return fail(ASTRAEA_E_STATE, "a queued step waits for a wider neighbour table: call astraea_synchronize");
        c->unsettled.clear();
    }
    uint32_t table[kScratchWords];
    float* com = reinterpret_cast<float*>(table);
    uint32_t* base = table + 3u * ASTRAEA_NOALLOC_MAX_SEGMENTS;
    uint32_t total = 0;
    for (uint32_t k = 0; k < S; ++k) {
        for (int a = 0; a < 3; ++a) com[3 * k + a] = (float)comPsu[3 * k + a];
        base[k] = total;
        total += c->counts[k];
    }
    uint32_t* dCom = c->scratchDev;
    uint32_t* dBase = c->scratchDev + 3u * ASTRAEA_NOALLOC_MAX_SEGMENTS;
    HIP_TRY(hipMemcpyAsync(dCom, com, sizeof(float) * 3 * S, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dBase, base, sizeof(uint32_t) * S, hipMemcpyHostToDevice, c->stream));
    launchPackRender(c->dc, c->stream, reinterpret_cast<const float*>(dCom), dBase,
                     reinterpret_cast<uint32_t*>(recordsDev), speedDev);
    TRY(checkLaunch(c));
    HIP_TRY(hipStreamSynchronize(c->stream));  // the host table is read before it goes out of scope
    return ASTRAEA_OK;
}

int astraea_set_tracers(astraea_ctx* c, uint32_t count, const uint32_t* particleId, const float* rest,
                        const uint8_t* ocean) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    if (count > 0 && (!particleId || !rest || !ocean)) return fail(ASTRAEA_E_INVALID, "null argument");
    if (count > 65535u) return fail(ASTRAEA_E_INVALID, "%u tracers: the record's anchor index is 16 bits", count);
    HIP_TRY(hipSetDevice(c->device));
    TRY(ensureRenderBuffers(c));
    uint32_t total = 0;
    for (uint32_t k = 0; k < c->numSeg; ++k) total += c->counts[k];
    std::vector<int32_t> of(total, -1);
    for (uint32_t t = 0; t < count; ++t) {
        if (particleId[t] >= total) return fail(ASTRAEA_E_INVALID, "tracer %u: particle id %u out of range", t, particleId[t]);
        if (of[particleId[t]] >= 0) return fail(ASTRAEA_E_INVALID, "particle %u carries two tracers", particleId[t]);
        of[particleId[t]] = (int32_t)t;
    }
    if (count > c->tracerCap) {
        // (re)allocated outside any step; the old tables are released with the context
        TRY(devAlloc(c, c->dTracerRest, count));
        TRY(devAlloc(c, c->dTracerOcean, count));
        TRY(devAlloc(c, c->dTracerOut, 10 * (size_t)count));
        c->tracerCap = count;
    }
    HIP_TRY(hipMemcpyAsync(c->dTracerOf, of.data(), sizeof(int32_t) * total, hipMemcpyHostToDevice, c->stream));
    if (count > 0) {
        HIP_TRY(hipMemcpyAsync(c->dTracerRest, rest, sizeof(float4) * count, hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(c->dTracerOcean, ocean, count, hipMemcpyHostToDevice, c->stream));
    }
    c->tracerCount = count;
    launchMarkTracers(c->dc, c->stream, c->dTracerOf, total);
    TRY(checkLaunch(c));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ASTRAEA_OK;
}

int astraea_extract_tracers(astraea_ctx* c, const double* livingComPsu, const double* sysComPsu, uint8_t* records) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    if (!livingComPsu || !sysComPsu || (c->tracerCount > 0 && !records)) return fail(ASTRAEA_E_INVALID, "null argument");
    if (c->tracerCount == 0) return ASTRAEA_OK;
    HIP_TRY(hipSetDevice(c->device));
    const uint32_t S = c->numSeg;
    uint32_t total = 0;
    for (uint32_t k = 0; k < S; ++k) total += c->counts[k];
    std::vector<float> com(6 * (size_t)S);
    for (uint32_t k = 0; k < 3 * S; ++k) {
        com[k] = (float)livingComPsu[k];
        com[3 * S + k] = (float)sysComPsu[k];
    }
    HIP_TRY(hipMemcpyAsync(c->dCom, com.data(), sizeof(float) * 6 * S, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemsetAsync(c->dTracerOut, 0, 40 * (size_t)c->tracerCount, c->stream));
    launchExtractTracers(c->dc, c->stream, c->dTracerOf, total, c->dTracerRest, c->dTracerOcean, c->tracerCount,
                         c->dCom, c->dCom + 3 * S, c->dTracerOut);
    TRY(checkLaunch(c));
    HIP_TRY(hipMemcpyAsync(records, c->dTracerOut, 40 * (size_t)c->tracerCount, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return ASTRAEA_OK;
}

int astraea_set_neighbour_stride(astraea_ctx* c, uint32_t stride) {
    if (!c) return fail(ASTRAEA_E_INVALID, "null context");
    if (stride == 0 || stride > kMaxNeighbourStride)
        return fail(ASTRAEA_E_INVALID, "neighbour stride %u outside [1, %u]", stride, kMaxNeighbourStride);
    stride = roundUp(stride, 8 * kBanks);  // kBanks lists per slot, in chunks of eight entries
    HIP_TRY(hipSetDevice(c->device));
    if (c->loaded) TRY(settleQueued(c));
    HIP_TRY(hipStreamSynchronize(c->stream));
    TRY(allocNeighbourList(c, stride));
    c->listValid = false;
    return ASTRAEA_OK;
}

int astraea_counters(astraea_ctx* c, uint64_t* out, int n) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    if (!out || n <= 0) return fail(ASTRAEA_E_INVALID, "null output");
    HIP_TRY(hipStreamSynchronize(c->stream));
    unsigned long long v[8];
    HIP_TRY(hipMemcpy(v, c->dc.cnt.v, sizeof(v), hipMemcpyDeviceToHost));
    unsigned long long q[4];
    HIP_TRY(hipMemcpy(q, c->dc.nb.q, sizeof(q), hipMemcpyDeviceToHost));
    uint64_t vals[10] = {v[0], v[1], 0, c->dc.nb.stride, 0, 0, 0, c->stepsRedone, q[0],
                         std::min<unsigned long long>(q[1], c->dc.nb.poolBytes)};
    {
        const uint32_t waves = c->slots / kWave;
        std::vector<uint32_t> ws(waves), wm(waves);
        HIP_TRY(hipMemcpy(ws.data(), c->dc.nb.waveSum, sizeof(uint32_t) * waves, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(wm.data(), c->dc.nb.waveMax, sizeof(uint32_t) * waves, hipMemcpyDeviceToHost));
        for (uint32_t w = 0; w < waves; ++w) {
            vals[2] += ws[w];
            vals[6] = std::max<uint64_t>(vals[6], wm[w]);
        }
    }
    if (n > 4) {
        std::vector<float4> a(c->slots), b(c->slots);
        HIP_TRY(hipMemcpy(a.data(), c->dc.d.p1a, sizeof(float4) * c->slots, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(b.data(), c->dc.d.rate2, sizeof(float4) * c->slots, hipMemcpyDeviceToHost));
        forEachParticle(c, [&](size_t, size_t s) {
            vals[4] += a[s].w >= 1.0f ? (uint64_t)a[s].w - 1 : 0;
            vals[5] += (uint64_t)b[s].w;
        });
    }
    for (int k = 0; k < n && k < 10; ++k) out[k] = vals[k];
    // interaction and forced-acceptance counts restart from zero
    HIP_TRY(hipMemset(c->dc.cnt.v, 0, 2 * sizeof(unsigned long long)));
    return ASTRAEA_OK;
}

int astraea_download_tree(astraea_ctx* c, uint32_t seg, uint64_t* keys, uint32_t* perm, int32_t* left, int32_t* right,
                          int32_t* pI, int32_t* pL, float* cm, float* mn, float* mx) {
    TRY(requireLoaded(c));
    TRY(settleQueued(c));
    if (seg >= c->numSeg) return fail(ASTRAEA_E_INVALID, "segment %u out of range", seg);
    if (!c->treeValid) return fail(ASTRAEA_E_STATE, "no tree: call astraea_sort_and_build_bvh");
    HIP_TRY(hipStreamSynchronize(c->stream));
    const uint32_t off = c->offs[seg], n = c->counts[seg], ni = n >= 2 ? n - 1 : 0;
    const int o = (int)off;
    auto toLocal = [&](int v) { return v < 0 ? ~((~v) - o) : v - o; };
    if (keys) HIP_TRY(hipMemcpy(keys, c->dc.sb.keyA + off, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
    if (perm) {
        HIP_TRY(hipMemcpy(perm, c->dc.sb.valA + off, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < n; ++k) perm[k] -= off;
    }
    if (ni > 0 && (left || right)) {
        std::vector<int2> ch(ni);
        HIP_TRY(hipMemcpy(ch.data(), c->dc.t.child + off, sizeof(int2) * ni, hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < ni; ++k) {
            if (left) left[k] = toLocal(ch[k].x);
            if (right) right[k] = toLocal(ch[k].y);
        }
    }
    if (ni > 0 && pI) {
        HIP_TRY(hipMemcpy(pI, c->dc.t.parentI + off, sizeof(int32_t) * ni, hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < ni; ++k) pI[k] = pI[k] < 0 ? -1 : pI[k] - o;
    }
    if (pL) {
        HIP_TRY(hipMemcpy(pL, c->dc.t.parentL + off, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < n; ++k) pL[k] = pL[k] < 0 ? -1 : pL[k] - o;
    }
    if (ni > 0 && cm) HIP_TRY(hipMemcpy(cm, c->dc.t.cm + off, sizeof(float4) * ni, hipMemcpyDeviceToHost));
    if (ni > 0 && mn) HIP_TRY(hipMemcpy(mn, c->dc.t.bmin + off, sizeof(float4) * ni, hipMemcpyDeviceToHost));
    if (ni > 0 && mx) HIP_TRY(hipMemcpy(mx, c->dc.t.bmax + off, sizeof(float4) * ni, hipMemcpyDeviceToHost));
    return ASTRAEA_OK;
}

// The (wavefronts per block, stack entries) pairs of the gravity walk and the neighbour build.
static constexpr uint32_t kWalkWaves[4] = {2u, 4u, 4u, 8u}, kWalkStack[4] = {1024u, 1024u, 512u, 512u};

static bool validWalk(uint32_t waves, uint32_t stack) {
    for (int k = 0; k < 4; ++k)
        if (waves == kWalkWaves[k] && stack == kWalkStack[k]) return true;
    return false;
}

static bool validTuning(const astraea_tuning& t) {
    return (t.pass1_round == 128u || t.pass1_round == 256u) && (t.pass2_round == 128u || t.pass2_round == 160u) &&
           (t.drift_round == 128u || t.drift_round == 160u || t.drift_round == 192u || t.drift_round == 256u) &&
           (t.sort_bits == 8u || t.sort_bits == 9u) && validWalk(t.gravity_waves, t.gravity_stack) &&
           validWalk(t.build_waves, t.build_stack);
}

int astraea_get_tuning(astraea_ctx* c, astraea_tuning* out) {
    if (!c || !out) return fail(ASTRAEA_E_INVALID, "null argument");
    out->pass1_round = c->dc.tune.pass1Round;
    out->pass2_round = c->dc.tune.pass2Round;
    out->drift_round = c->dc.tune.driftRound;
    out->sort_bits = c->dc.tune.sortBits;
    out->gravity_waves = c->dc.tune.gravityWaves;
    out->gravity_stack = c->dc.tune.gravityStack;
    out->build_waves = c->dc.tune.buildWaves;
    out->build_stack = c->dc.tune.buildStack;
    return ASTRAEA_OK;
}

int astraea_set_tuning(astraea_ctx* c, const astraea_tuning* t) {
    if (!c || !t) return fail(ASTRAEA_E_INVALID, "null argument");
    if (!validTuning(*t))
        return fail(ASTRAEA_E_INVALID,
                    "tuning (%u, %u, %u, %u, gravity %u x %u, build %u x %u) outside pass 1 {128, 256}, pass 2 "
                    "{128, 160}, drift {128, 160, 192, 256}, sort bits {8, 9}, (waves, stack) {(2, 1024), (4, 1024), "
                    "(4, 512), (8, 512)}",
                    t->pass1_round, t->pass2_round, t->drift_round, t->sort_bits, t->gravity_waves, t->gravity_stack,
                    t->build_waves, t->build_stack);
    c->dc.tune.pass1Round = t->pass1_round;
    c->dc.tune.pass2Round = t->pass2_round;
    c->dc.tune.driftRound = t->drift_round;
    c->dc.tune.sortBits = t->sort_bits;
    c->dc.tune.gravityWaves = t->gravity_waves;
    c->dc.tune.gravityStack = t->gravity_stack;
    c->dc.tune.buildWaves = t->build_waves;
    c->dc.tune.buildStack = t->build_stack;
    freeGraph(c);  // a captured step holds the old kernels
    return ASTRAEA_OK;
}

extern "C++" {
// Milliseconds of the work `launch` enqueues on the context stream, timed with events.
template <class F>
static int timeLaunch(astraea_ctx* c, F&& launch, double* ms) {
    hipEvent_t e[2] = {nullptr, nullptr};
    int rc = ASTRAEA_OK;
    for (auto& x : e)
        if (rc == ASTRAEA_OK && hipEventCreate(&x) != hipSuccess) rc = fail(ASTRAEA_E_HIP, "hipEventCreate failed");
    if (rc == ASTRAEA_OK && hipEventRecord(e[0], c->stream) != hipSuccess) rc = fail(ASTRAEA_E_HIP, "hipEventRecord failed");
    if (rc == ASTRAEA_OK) {
        launch();
        rc = checkLaunch(c);
    }
    if (rc == ASTRAEA_OK && hipEventRecord(e[1], c->stream) != hipSuccess) rc = fail(ASTRAEA_E_HIP, "hipEventRecord failed");
    if (rc == ASTRAEA_OK && hipEventSynchronize(e[1]) != hipSuccess) rc = fail(ASTRAEA_E_HIP, "hipEventSynchronize failed");
    float t = 0.0f;
    if (rc == ASTRAEA_OK && hipEventElapsedTime(&t, e[0], e[1]) != hipSuccess) rc = fail(ASTRAEA_E_HIP, "hipEventElapsedTime failed");
    for (auto& x : e)
        if (x) (void)hipEventDestroy(x);
    *ms = t;
    return rc;
}

// The (waves, stack) variants of the gravity walk and the neighbour build, timed by launching them
// again on the settled state (they rewrite what the last step computed, bit for bit): eight runs in
// palindromic order.  A gravity variant that force-accepts a node or a build variant whose stack
// fills is ruled out (time -1).  The table is then built again with the build variant chosen.
static int tuneWalks(astraea_ctx* c, double gms[4], double bms[4], int* gBest, int* bBest) {
    static constexpr int order4[8] = {0, 1, 2, 3, 3, 2, 1, 0};
    bool gOk[4] = {true, true, true, true}, bOk[4] = {true, true, true, true};
    for (int k = 0; k < 4; ++k) gms[k] = bms[k] = 0.0;
    unsigned long long saved[2];  // the interaction and forced-acceptance counts of the steps
    HIP_TRY(hipMemcpy(saved, c->dc.cnt.v, sizeof(saved), hipMemcpyDeviceToHost));
    const Tuning keep = c->dc.tune;
    int rc = ASTRAEA_OK;
    // every variant once, untimed, so that no timed run carries a kernel's first launch
    for (int k = 0; k < 4 && rc == ASTRAEA_OK; ++k) {
        c->dc.tune.gravityWaves = kWalkWaves[k];
        c->dc.tune.gravityStack = kWalkStack[k];
        double t = 0.0;
        rc = timeLaunch(c, [&] { launchGravity(c->dc, c->stream); }, &t);
    }
    for (int r = 0; r < 8 && rc == ASTRAEA_OK; ++r) {
        const int k = order4[r];
        c->dc.tune.gravityWaves = kWalkWaves[k];
        c->dc.tune.gravityStack = kWalkStack[k];
        double t = 0.0;
        rc = timeLaunch(c, [&] { launchGravity(c->dc, c->stream); }, &t);
        unsigned long long v[8];
        if (rc == ASTRAEA_OK && hipMemcpy(v, c->dc.cnt.v, sizeof(v), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(ASTRAEA_E_HIP, "hipMemcpy failed");
        if (rc != ASTRAEA_OK) break;
        if (v[1] != saved[1]) gOk[k] = false;
        gms[k] += 0.5 * t;
    }
    for (int k = 0; k < 4 && rc == ASTRAEA_OK; ++k) {
        c->dc.tune.buildWaves = kWalkWaves[k];
        c->dc.tune.buildStack = kWalkStack[k];
        double t = 0.0;
        rc = timeLaunch(c, [&] { launchNeighbours(c->dc, c->stream); }, &t);
        unsigned long long v[8];
        if (rc == ASTRAEA_OK && hipMemcpy(v, c->dc.cnt.v, sizeof(v), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(ASTRAEA_E_HIP, "hipMemcpy failed");
        if (rc == ASTRAEA_OK && v[kCntHold] != 0ull) {
            HIP_TRY(hipMemset(c->dc.cnt.v + kCntHold, 0, 2 * sizeof(unsigned long long)));
            HIP_TRY(hipMemset(c->dc.cnt.v + 5, 0, 3 * sizeof(unsigned long long)));
        }
    }
    for (int r = 0; r < 8 && rc == ASTRAEA_OK; ++r) {
        const int k = order4[r];
        c->dc.tune.buildWaves = kWalkWaves[k];
        c->dc.tune.buildStack = kWalkStack[k];
        double t = 0.0;
        rc = timeLaunch(c, [&] { launchNeighbours(c->dc, c->stream); }, &t);
        unsigned long long v[8];
        if (rc == ASTRAEA_OK && hipMemcpy(v, c->dc.cnt.v, sizeof(v), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(ASTRAEA_E_HIP, "hipMemcpy failed");
        if (rc != ASTRAEA_OK) break;
        if (v[6] != 0ull) bOk[k] = false;
        bms[k] += 0.5 * t;
        // a held build (a full stack, or a spill pool the variant found too small) leaves the
        // table incomplete: clear the flags; the chosen variant builds it again below
        if (v[kCntHold] != 0ull) {
            HIP_TRY(hipMemset(c->dc.cnt.v + kCntHold, 0, 2 * sizeof(unsigned long long)));
            HIP_TRY(hipMemset(c->dc.cnt.v + 5, 0, 3 * sizeof(unsigned long long)));
        }
    }
    // restore the step's counts (the gravity launches above added interactions)
    if (rc == ASTRAEA_OK && hipMemcpy(c->dc.cnt.v, saved, sizeof(saved), hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(ASTRAEA_E_HIP, "hipMemcpy failed");
    c->dc.tune = keep;
    if (rc != ASTRAEA_OK) return rc;
    int g = -1, b = -1;
    for (int k = 0; k < 4; ++k) {
        if (gOk[k] && (g < 0 || gms[k] < gms[g])) g = k;
        if (bOk[k] && (b < 0 || bms[k] < bms[b])) b = k;
        if (!gOk[k]) gms[k] = -1.0;
        if (!bOk[k]) bms[k] = -1.0;
    }
    *gBest = g < 0 ? 1 : g;  // (4, 1024) when every variant filled its stack
    *bBest = b < 0 ? 1 : b;
    c->dc.tune.buildWaves = kWalkWaves[*bBest];
    c->dc.tune.buildStack = kWalkStack[*bBest];
    launchNeighbours(c->dc, c->stream);
    TRY(checkLaunch(c));
    rc = fitNeighbourTable(c);
    c->dc.tune = keep;
    return rc;
}
}  // extern "C++"

int astraea_autotune(astraea_ctx* c, float maxDt, uint32_t stepsPerVariant, astraea_tuning* chosen, double* ms) {
    if (!c) return fail(ASTRAEA_E_INVALID, "null context");
    if (stepsPerVariant == 0) return fail(ASTRAEA_E_INVALID, "steps per variant must be > 0");
    static constexpr uint32_t p1[2] = {128u, 256u}, p2[2] = {128u, 160u}, dr[4] = {128u, 160u, 192u, 256u},
                              sb[2] = {8u, 9u};
    // eight runs in palindromic orders: every variant's runs sit on average at the middle of the
    // sequence, so a steady drift of the step's cost as the state evolves cancels out
    static constexpr int order2[8] = {0, 1, 1, 0, 0, 1, 1, 0}, order4[8] = {0, 1, 2, 3, 3, 2, 1, 0};
    const bool wasTiming = c->timing;
    astraea_tuning before;
    TRY(astraea_get_tuning(c, &before));
    double t1[2] = {0.0, 0.0}, t2[2] = {0.0, 0.0}, td[4] = {0.0, 0.0, 0.0, 0.0}, ts[2] = {0.0, 0.0};
    // one untimed step with each variant first, so that no timed step carries a kernel's first
    // launch or the timing's set-up
    int rc = ASTRAEA_OK;
    for (int r = 0; r < 4 && rc == ASTRAEA_OK; ++r) {
        astraea_tuning t = before;
        t.pass1_round = p1[r % 2];
        t.pass2_round = p2[r % 2];
        t.drift_round = dr[r];
        t.sort_bits = sb[r % 2];
        rc = astraea_set_tuning(c, &t);
        if (rc == ASTRAEA_OK) rc = astraea_step(c, maxDt, 1, 0);
        if (rc == ASTRAEA_OK) rc = settle(c);
    }
    if (rc == ASTRAEA_OK) rc = astraea_enable_stage_timing(c, 1);
    for (int r = 0; r < 8 && rc == ASTRAEA_OK; ++r) {
        astraea_tuning t = before;
        t.pass1_round = p1[order2[r]];
        t.pass2_round = p2[order2[r]];
        t.drift_round = dr[order4[r]];
        t.sort_bits = sb[order2[r]];
        rc = astraea_set_tuning(c, &t);
        if (rc == ASTRAEA_OK) rc = astraea_step(c, maxDt, stepsPerVariant, 0);
        if (rc == ASTRAEA_OK) rc = settle(c);
        double stage[ASTRAEA_STAGE_COUNT_];
        uint64_t n[ASTRAEA_STAGE_COUNT_];
        if (rc == ASTRAEA_OK) rc = astraea_stage_times(c, stage, n);
        if (rc != ASTRAEA_OK) break;
        const auto mean = [&](int k) { return n[k] ? stage[k] / (double)n[k] : 0.0; };
        t1[order2[r]] += 0.25 * mean(ASTRAEA_STAGE_PASS1);
        t2[order2[r]] += 0.25 * mean(ASTRAEA_STAGE_PASS2);
        td[order4[r]] += 0.5 * mean(ASTRAEA_STAGE_DRIFT);
        ts[order2[r]] += 0.25 * mean(ASTRAEA_STAGE_SORT);
    }
    const int rt = astraea_enable_stage_timing(c, wasTiming ? 1 : 0);
    double gms[4] = {0.0, 0.0, 0.0, 0.0}, bms[4] = {0.0, 0.0, 0.0, 0.0};
    int gBest = 1, bBest = 1;
    if (rc == ASTRAEA_OK) rc = tuneWalks(c, gms, bms, &gBest, &bBest);
    if (rc != ASTRAEA_OK) {
        (void)astraea_set_tuning(c, &before);
        return rc;
    }
    TRY(rt);
    int bd = 0;
    for (int k = 1; k < 4; ++k)
        if (td[k] < td[bd]) bd = k;
    const astraea_tuning best = {t1[0] < t1[1] ? p1[0] : p1[1], t2[0] < t2[1] ? p2[0] : p2[1], dr[bd],
                                 ts[0] < ts[1] ? sb[0] : sb[1], kWalkWaves[gBest], kWalkStack[gBest],
                                 kWalkWaves[bBest], kWalkStack[bBest]};
    TRY(astraea_set_tuning(c, &best));
    if (chosen) *chosen = best;
    if (ms) {
        const double all[18] = {t1[0], t1[1], t2[0], t2[1], td[0], td[1], td[2], td[3], ts[0], ts[1],
                                gms[0], gms[1], gms[2], gms[3], bms[0], bms[1], bms[2], bms[3]};
        for (int k = 0; k < 18; ++k) ms[k] = all[k];
    }
    return ASTRAEA_OK;
}

int astraea_enable_stage_timing(astraea_ctx* c, int enable) {
    if (!c) return fail(ASTRAEA_E_INVALID, "null context");
    if (enable && c->evPool.empty()) {
        c->evPool.resize(astraea_ctx::kPool, nullptr);
        for (auto& e : c->evPool) HIP_TRY(hipEventCreate(&e));
    }
    collectTimes(c);
    c->timing = enable != 0;
    for (int k = 0; k < ASTRAEA_STAGE_COUNT_; ++k) { c->stageMs[k] = 0.0; c->stageN[k] = 0; }
    return ASTRAEA_OK;
}

int astraea_stage_times(astraea_ctx* c, double* ms, uint64_t* samples) {
    if (!c) return fail(ASTRAEA_E_INVALID, "null context");
    collectTimes(c);
    for (int k = 0; k < ASTRAEA_STAGE_COUNT_; ++k) {
        if (ms) ms[k] = c->stageMs[k];
        if (samples) samples[k] = c->stageN[k];
        c->stageMs[k] = 0.0;
        c->stageN[k] = 0;
    }
    return ASTRAEA_OK;
}

}  // extern "C"
