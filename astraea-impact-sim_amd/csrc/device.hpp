// device.hpp — device data layout and shared helpers of the gfx950 kernels.
//
// Layout in HBM (one context):
//   * particle slots are Morton-ordered inside each segment (scenario); segment s
//     owns slots [segOff[s], segOff[s] + segPad[s]) with segOff a multiple of
//     ASTRAEA_SEGMENT_ALIGN (= the 256-thread block), so a block and a wavefront
//     never straddle two scenarios.  The first segN[s] slots are particles, the
//     rest are inactive padding.
//   * persistent state, ping-ponged through the sort's permutation, packed as
//     float4 so one lane moves one 16-byte record:
//       pos   (x, y, z, m)          vel (vx, vy, vz, rho)      hu (h, h0, u, D)
//       str0  (sxx, syy, szz, sxy)  str1 (sxz, syz, D^(1/3), eps_fail)
//       meta  (material | body << 8 | tracer << 16)            oid (input index)
//     plus FP64 x/y/z and vx/vy/vz in the FP64-accumulator mode.
//   * derived per-step arrays in the current order (not permuted).
//   * tree: internal node of local index i of segment s lives at slot
//     segOff[s] + i; a child c < 0 is the particle slot ~c.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/astraea_hip.h"

namespace astraea_hip {

constexpr int kBlock = 256;                 // threads per block for every kernel
constexpr int kWave = 64;                   // CDNA4 wavefront
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kStack = ASTRAEA_STACK_DEPTH;
constexpr float kPiF = 3.14159265358979323846f;
constexpr float kGf = 66.7430f;             // G in PSU, engine/units.hpp:28
constexpr uint64_t kPadKey = 0x7fffffffffffffffULL;

struct Persist {
    float4* pos;
    float4* vel;
    float4* hu;
    float4* str0;
    float4* str1;
    uint32_t* meta;
    uint32_t* oid;
    double* p64;   // 3 arrays of cap doubles: x | y | z  (FP64 accumulator mode)
    double* v64;   // 3 arrays: vx | vy | vz
};

struct Derived {
    float4* posNew;   // drift output (x', y', z', m) in the pre-sort order
    double* blockBounds; // [6 per block] min xyz / max xyz of the block's positions (FP32 values,
                         // or the FP64 positions' with FP64 accumulators)
    float4* grav;     // (gx, gy, gz, phi)
    float4* rate;     // (ax, ay, az, drho/dt): total acceleration after pass 2
    float4* rate2;    // (du/dt, dD^(1/3)/dt, radPowerW, 0)
    float4* dS0;      // (dsxx, dsyy, dszz, dsxy)
    float2* dS1;      // (dsxz, dsyz)
    float4* eos;      // (P, cs, T, 0) of the last EOS evaluation
    uint8_t* phase;
    float4* p1a;      // (Shepard C, f_surf, div v, neighbour count incl. self)
    float4* p1b;      // (curl v x, y, z, Balsara)
    float4* nb;       // pass-2 neighbour record (h', P_eff/rho^2, max(cs,0.05), Balsara)
    float4* sr0;      // S/rho^2 (xx, yy, zz, xy) for pass 2
    float2* sr1;      // S/rho^2 (xz, yz)
};

struct Tree {
    int2* child;          // [slot of internal node] (left, right)
    int32_t* parentI;     // parent of internal node
    int32_t* parentL;     // parent of leaf slot
    uint32_t* visited;    // bottom-up arrival counters (nodes whose leaves span two moment chunks)
    int32_t* other;       // slot of the far end of an internal node's leaf range (the near end is the node's own index)
    float4* cm;           // (com x, y, z, mass)
    float4* bmin;         // (min x, y, z, extent = max side)
    float4* bmax;         // (max x, y, z, max support bound of the subtree)
};

// Banks of the neighbour table (see WaveTable).
constexpr uint32_t kBanks = 4;

// Where a wavefront's part of the neighbour table lives.  Its union: the slots that are a
// neighbour of at least one of its particles; the q-th slot the walk found goes to bank q % kBanks
// at position q / kBanks, bank b at uni + b ucap / kBanks.  A particle's neighbours come in runs of
// slots the walk found one after another, so dealing the union out to the banks like cards gives
// every particle about the same number of entries in each bank, and the rounds of a pass (a bank
// at a time, staged in LDS) keep the lanes equally busy; in walk order a particle's neighbours
// would crowd into a few rounds.  An entry is a 16-bit position in its bank, and a particle's
// entries of one bank ascend.  Each slot's `stride` entries are kBanks lists of stride / kBanks;
// entry k of bank b of lane l is half-word (k & 7) of chunk 64 (b stride / (8 kBanks) + (k >> 3)) + l
// of `list`: a lane's entries come in 16-byte chunks of eight, chunk r of the 64 lanes is one
// contiguous kilobyte.
struct WaveTable {
    uint16_t* list;
    uint32_t* uni;
    uint32_t stride;      // entries per slot, a multiple of 8 kBanks
    uint32_t ucap;        // union ids, a multiple of kBanks (<= 65536)
};

struct Neighbours {
    uint2* count;         // entries of each slot per bank (four u16, bank b in half b & 1 of word b >> 1)
    uint32_t* waveSum;    // entries wanted by each wavefront's particles (uncapped)
    uint32_t* waveMax;    // largest entry count one particle of the wavefront wanted
    uint32_t* waveNeed;   // table width the wavefront needs: kBanks x its fullest bank, rounded up to 32
    uint32_t* unionLen;   // per wavefront: length of its union (uncapped)
    WaveTable* tab;       // per wavefront: where its table lives
    uint16_t* list;       // regular region: wavefront w at list + 64 w stride
    uint32_t* unionIds;   // regular region: wavefront w at unionIds + w unionCap
    uint32_t stride;      // regular entries per slot
    uint32_t unionCap;    // regular union ids per wavefront
    uint32_t* spill;      // wavefronts the regular region cannot hold, in the order they were found
    unsigned long long* q;  // [4]: 0 wavefronts listed, 1 pool bytes reserved (also past the end),
                            //      2 and 3 unused
    uint8_t* pool;        // the spill pool
    unsigned long long poolBytes;
};

__device__ __forceinline__ uint32_t bankCount(uint2 c, uint32_t b) {
    return ((b < 2u ? c.x : c.y) >> (16u * (b & 1u))) & 0xffffu;
}

// Orders the LDS accesses of a wavefront's lanes around this point: the lanes of one wavefront
// run in lockstep and LDS serves a wavefront's accesses in order, so only the compiler must be
// kept from moving accesses across.
__device__ __forceinline__ void waveSync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Walks the neighbours of slot i.  Every lane of the wavefront must call it together (inactive
// slots have no entries).  Bank by bank, the wave's union is staged through `buf` (S records per
// round): the round's slot ids are loaded together, then load(slot, rec) fills the records with
// coalesced loads and finish(rec) derives what a record needs from them; each lane consumes its
// entries that fall in the round, pair(rec) in table order, reading the next entry's record from
// LDS before it runs pair on the current one.  Past the end of a bank a round stages copies of
// its last slot (never read).  S is a multiple of 32: a round is ceil(S / 64) wavefront loads, and
// in the last one of a round of 160 only lanes 0-31 store their record.
//
// (Keeping the next round's records in flight while a round is evaluated does not pay here: the
// wait for an entry chunk inside the pair loop is a wait for every older load, and the records
// held across the round cost a wavefront per SIMD.)
//
// The records sit in LDS quad-major: 16-byte quad q of record k at quad q * S + k of the buffer
// (the same S * sizeof(Rec) bytes).  Lanes read records at scattered positions; with records
// stored whole, a record of 2 or 4 quads would start on only 8 or 4 of the 16 four-bank groups,
// while quad-major rows spread the lanes over all 16.
template <int S, class Rec>
__device__ __forceinline__ Rec ldsRecord(const float4* q, uint32_t k) {
    constexpr int NQ = (int)(sizeof(Rec) / 16);
    float4 t[NQ];
#pragma unroll
    for (int x = 0; x < NQ; ++x) t[x] = q[x * S + k];
    Rec r;
    __builtin_memcpy(&r, t, sizeof(Rec));
    return r;
}

// Native 4- and 2-float vectors for the records a neighbour pass stages: the HIP vector types
// copy through private memory when they are assigned, which keeps a record that lives across a
// loop iteration out of registers; these are plain register values.
typedef float Vec4 __attribute__((ext_vector_type(4)));
typedef float Vec2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ Vec4 loadVec(const float4* a, uint32_t j) { return reinterpret_cast<const Vec4*>(a)[j]; }
__device__ __forceinline__ Vec2 loadVec(const float2* a, uint32_t j) { return reinterpret_cast<const Vec2*>(a)[j]; }
__device__ __forceinline__ float4 toFloat4(Vec4 v) { return make_float4(v.x, v.y, v.z, v.w); }
__device__ __forceinline__ float2 toFloat2(Vec2 v) { return make_float2(v.x, v.y); }

// A pointer read from device memory (a WaveTable's) is generic to the compiler, which then
// issues flat loads: they count against both the vector-memory and the LDS wait counters, so
// every wait for one also waits for LDS.  The table lives in global memory; this says so.
template <class T>
using GlobalPtr = const T __attribute__((address_space(1)))*;
template <class T>
__device__ __forceinline__ GlobalPtr<T> globalView(const T* p) {
    return (GlobalPtr<T>)p;
}

template <int S, class Rec, class Load, class Finish, class Pair>
__device__ __forceinline__ void forEachNeighbour(const Neighbours& nb, uint32_t i, Rec* buf, Load&& load,
                                                 Finish&& finish, Pair&& pair) {
    static_assert(S % 32 == 0, "a round is whole half-wavefront loads");
    static_assert(sizeof(Rec) % 16 == 0 && alignof(Rec) >= 16, "records are whole 16-byte quads");
    constexpr int NQ = (int)(sizeof(Rec) / 16);
    float4* quads = reinterpret_cast<float4*>(buf);
    constexpr int R = (S + 63) / 64;  // wavefront loads per round; the last may fill half the lanes
    constexpr uint32_t kPast = 0xffffffffu;  // an entry position beyond every round
    const uint32_t w = i >> 6, lane = i & 63u;
    const WaveTable t = nb.tab[w];
    const uint32_t total = __builtin_amdgcn_readfirstlane(min(nb.unionLen[w], t.ucap));
    const uint32_t bankCap = t.ucap / kBanks;
    const uint2 counts = nb.count[i];
    const uint32_t perBank = t.stride / (8u * kBanks);
    const uint32_t lastChunk = perBank - 1u;
    // entry chunks as native vectors: a HIP uint4 read through a global pointer becomes a copy
    // through private memory
    typedef uint32_t Chunk __attribute__((ext_vector_type(4)));
    const GlobalPtr<Chunk> chunks = globalView(reinterpret_cast<const Chunk*>(t.list)) + lane;
#pragma unroll 1
    for (uint32_t bank = 0; bank < kBanks; ++bank) {
        const uint32_t len = (total + (kBanks - 1u) - bank) / kBanks;
        if (len == 0u) break;  // banks fill in turn: the later ones are empty too
        const GlobalPtr<uint32_t> U = globalView(t.uni) + bank * bankCap;
        const uint32_t cnt = bankCount(counts, bank);
        // chunks are loaded one ahead, unconditionally (clamped to the bank's last chunk: entries
        // past cnt are never used), so that no load result has to be merged with an older value
        const GlobalPtr<Chunk> ch = chunks + 64ull * (bank * perBank);
        Chunk cur = ch[0];
        Chunk nxt = ch[64ull * min(1u, lastChunk)];
        uint32_t k = 0;
        uint32_t e = cnt > 0u ? (cur.x & 0xffffu) : kPast;  // position of entry k; entries ascend, so e >= s0
        // moves to entry k + 1 and returns its position (kPast after the last)
        const auto advance = [&]() -> uint32_t {
            ++k;
            if ((k & 7u) == 0u) {
                cur = nxt;
                nxt = ch[64ull * min((k >> 3) + 1u, lastChunk)];
            } else if ((k & 1u) == 0u) {
                cur.x = cur.y;
                cur.y = cur.z;
                cur.z = cur.w;
            }
            return k < cnt ? ((k & 1u) ? (cur.x >> 16) : (cur.x & 0xffffu)) : kPast;
        };
        for (uint32_t s0 = 0; s0 < len; s0 += S) {
            const uint32_t s1 = min(s0 + (uint32_t)S, len);
            uint32_t id[R];
#pragma unroll
            for (int r = 0; r < R; ++r) id[r] = U[min(s0 + (uint32_t)(64 * r) + lane, len - 1u)];
            waveSync();
#pragma unroll
            for (int r = 0; r < R; ++r) {
                Rec rec;
                load(id[r], rec);
                finish(rec);
                float4 q[NQ];
                __builtin_memcpy(q, &rec, sizeof(Rec));
#pragma unroll
                for (int x = 0; x < NQ; ++x)
                    if (64 * r + (int)lane < S) quads[x * S + 64 * r + lane] = q[x];
            }
            waveSync();
            // two records alternate (a holds entry k when the loop body starts), so that a record
            // read ahead from LDS never has to be copied into the other's registers
            Rec a = ldsRecord<S, Rec>(quads, min(e - s0, (uint32_t)S - 1u));
            while (e < s1) {
                uint32_t en = advance();
                const Rec b = ldsRecord<S, Rec>(quads, min(en - s0, (uint32_t)S - 1u));
                pair(a);
                e = en;
                if (e >= s1) break;
                en = advance();
                a = ldsRecord<S, Rec>(quads, min(en - s0, (uint32_t)S - 1u));
                pair(b);
                e = en;
            }
        }
    }
    waveSync();
}

struct Segments {
    uint32_t count;       // number of segments
    uint32_t* off;        // [count+1] padded slot offsets
    uint32_t* n;          // [count] particles
    uint32_t* slotSeg;    // [slots] segment of every slot
};

// Per-segment device scalars.
struct SegScalars {
    double* bounds;       // [6 per segment] min xyz / max xyz of the positions being sorted (the Morton
                          // box: FP32 values, or FP64 with FP64 accumulators)
    uint32_t* dtRawBits;  // float bits of min dt_i (atomicMin target)
    float* dt;            // dt of the current step
    float* dtPrev;        // previous dt (0 before the first step)
    double* radSum;       // sum of radPowerW after the last pass 2 (W)
    double* radPrev;      // radSum of the previous step (W)
    double* erad;         // radiated energy integral (PSU energy)
    double* simTime;      // sum of dt (PSU time)
    uint32_t* living;     // any tracer anchor with D < 0.8 (sph_gravity.cu:478-483)
};

struct Counters {
    unsigned long long* v;  // [8]: 0 interactions, 1 stack overflows (gravity), 2 hold flag (see
                            // stepHeld), 3 nbr overflow flag, 4 steps completed (kStepEnd),
                            // 5 largest union a wavefront wanted (sticky max), 6 nbr stack overflow,
                            // 7 table width wanted: kBanks x the fullest bank of a slot (sticky max)
};
constexpr int kCntHold = 2;
constexpr int kCntSteps = 4;

struct Params {
    float theta2;
    uint32_t theta2Pow2;  // theta^2 is a power of two: multiplying by it is exact
    float softFactor;
    float cfl;
    float gammaHHe;
    float coolMinT;
    uint32_t strengthMask;
    uint32_t group;       // gravity opening group: 64 (a wavefront) or 32 (lanes 0-31 of one)
};

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ int laneId() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ int uniformInt(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t uniformU32(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ float readLaneF(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Logical block index that keeps consecutive blocks on one XCD: the dispatcher
// places block b on XCD b % 8, so XCD x receives a contiguous range of logical
// blocks and its private L2 holds one region of the Morton-ordered domain.
__device__ __forceinline__ uint32_t xcdBlock(uint32_t b, uint32_t nb) {
    const uint32_t x = b & 7u, i = b >> 3, per = nb >> 3, rem = nb & 7u;
    return x * per + (x < rem ? x : rem) + i;
}

// Squared distance, ((dx*dx) + dy*dy) + dz*dz as two FMAs: the operation order every
// neighbour and opening decision uses (reproduced by the oracle's FP32 replica).
// max(x, c) for c > 0 (or x, c both >= 0), as an integer max of the bit patterns: the bits of
// non-negative floats order like the floats, and a negative x loses to c either way.  fmaxf on a
// value loaded from memory costs an extra canonicalising instruction first; this does not (NaN
// inputs are the only difference).
__device__ __forceinline__ float maxPos(float x, float c) {
    return __int_as_float(max(__float_as_int(x), __float_as_int(c)));
}

__device__ __forceinline__ float dist2f(float dx, float dy, float dz) {
#pragma clang fp contract(off)
    float t = dx * dx;
    t = __builtin_fmaf(dy, dy, t);
    t = __builtin_fmaf(dz, dz, t);
    return t;
}

__device__ __forceinline__ float boxDist2f(float px, float py, float pz, float4 mn, float4 mx) {
#pragma clang fp contract(off)
    const float dx = fmaxf(0.0f, fmaxf(mn.x - px, px - mx.x));
    const float dy = fmaxf(0.0f, fmaxf(mn.y - py, py - mx.y));
    const float dz = fmaxf(0.0f, fmaxf(mn.z - pz, pz - mx.z));
    return dist2f(dx, dy, dz);
}

// Largest support 2*max(h',1e-6) a particle can reach in pass 1's update
// h' = clamp(h * clamp(cbrt(56/n), 0.92, 1.08), 0.35 h0, 6 h0) (sph_gravity.cu:785-787):
// rounding is monotone, so h' <= max(1.08 h, 0.35 h0) holds in FP32 too.
__device__ __forceinline__ float supportBound(float h, float h0) {
#pragma clang fp contract(off)
    return 2.0f * fmaxf(fmaxf(fmaxf(h, 1.0e-6f) * 1.08f, 0.35f * fmaxf(h0, 1.0e-6f)), 1.0e-6f);
}

// Ordered-uint encoding of floats for atomic min/max (sph_gravity.cu:140-149).
__device__ __forceinline__ uint32_t floatToOrdered(float f) {
    const uint32_t u = __float_as_uint(f);
    return u ^ ((uint32_t)(-(int32_t)(u >> 31)) | 0x80000000u);
}
__device__ __forceinline__ float orderedToFloat(uint32_t u) {
    return __uint_as_float(u ^ (((u >> 31) - 1u) | 0x80000000u));
}

// Wendland C2 kernel and gradient coefficient, support H (sph_gravity.cu:199-220).
__device__ __forceinline__ float wendland(float r, float H) {
    if (r >= H || H <= 0.0f) return 0.0f;
    const float q = r / H, o = 1.0f - q, o2 = o * o;
    return 21.0f / (2.0f * kPiF * H * H * H) * o2 * o2 * (1.0f + 4.0f * q);
}
__device__ __forceinline__ float wendlandGrad(float r, float H) {
    if (r >= H || H <= 0.0f) return 0.0f;
    const float q = r / H, o = 1.0f - q, H2 = H * H;
    return -(210.0f / (kPiF * H2 * H2 * H)) * o * o * o;
}

// The same kernels for the pair loops, from the per-particle constants of a support H:
// invH = 1/H, wNorm = 21/(2 pi H^3), gNorm = -210/(pi H^5).  They use the hardware
// reciprocal and square root (about 1 ulp), and vanish for r >= H as above.
struct Support {
    float invH, wNorm, gNorm;
};
__device__ __forceinline__ float fastRcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fastSqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ Support supportOf(float H) {
    const float iv = fastRcp(H), iv2 = iv * iv;
    return Support{iv, (21.0f / (2.0f * kPiF)) * iv2 * iv, (-210.0f / kPiF) * iv2 * iv2 * iv};
}
__device__ __forceinline__ float wendlandS(float r, const Support& s) {
    const float q = r * s.invH, o = fmaxf(0.0f, 1.0f - q), o2 = o * o;
    return s.wNorm * (o2 * o2) * (1.0f + 4.0f * q);
}
__device__ __forceinline__ float wendlandGradS(float r, const Support& s) {
    const float o = fmaxf(0.0f, 1.0f - r * s.invH);
    return s.gNorm * (o * o * o);
}

// Strength table (engine/sph_gravity.cu:35-73) gated by the params' strength mask.
struct Strength {
    float mu, tMelt, bulk, y0;
};
__device__ __forceinline__ Strength strengthOf(int mat, uint32_t mask) {
    const int mi = (mat >= 0 && mat <= 4) ? mat : 1;
    Strength s{0.0f, 0.0f, 0.0f, 0.0f};
    if (!((mask >> mi) & 1u)) return s;
    switch (mi) {
        case 0: s = {7.70e10f / 1.0e12f, 1811.0f, 1.28e11f / 1.0e12f, 6.80e8f / 1.0e12f}; break;
        case 1: s = {2.20e10f / 1.0e12f, 1600.0f, 1.80e10f / 1.0e12f, 3.50e9f / 1.0e12f}; break;
        case 2: s = {3.50e9f / 1.0e12f, 273.15f, 2.18e9f / 1.0e12f, 1.00e7f / 1.0e12f}; break;
        case 4: s = {8.00e9f / 1.0e12f, 1500.0f, 1.20e10f / 1.0e12f, 5.00e7f / 1.0e12f}; break;
        default: break;
    }
    return s;
}

// Per-material reference energy c_v * 300 K in PSU (sph_gravity.cu:27-33).
__device__ __forceinline__ float u0Ref(int mat) {
    switch ((mat >= 0 && mat <= 4) ? mat : 1) {
        case 0: return (float)((449.0 * 300.0) / 1.0e6);
        case 2: return (float)((4184.0 * 300.0) / 1.0e6);
        case 3: return (float)((8260.0 * 300.0) / 1.0e6);
        case 4: return (float)((800.0 * 300.0) / 1.0e6);
        default: return (float)((790.0 * 300.0) / 1.0e6);
    }
}

__device__ __forceinline__ float waveMinF(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float waveMaxF(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double waveSumD(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double waveMinD(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double waveMaxD(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ uint32_t waveMaxU(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t w = (uint32_t)__shfl_xor((int)v, o, 64);
        v = v > w ? v : w;
    }
    return v;
}

}  // namespace astraea_hip
