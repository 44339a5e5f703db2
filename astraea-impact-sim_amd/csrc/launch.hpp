// launch.hpp — host-side launchers of the gfx950 kernels (one translation unit per stage).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device.hpp"

namespace astraea_hip {

// Elements per tile of the radix sort (one 256-thread block, eight per thread).
constexpr uint32_t kSortTile = 2048;
constexpr uint32_t kSortMaxDigits = 512;  // 9-bit digits
// Passes of `bits`-bit digits over the 63 key bits.
constexpr int sortPassesOf(uint32_t bits) { return (int)((63u + bits - 1u) / bits); }

struct SortBuffers {
    uint64_t* keyA;
    uint64_t* keyB;
    uint32_t* valA;
    uint32_t* valB;
    uint32_t* counts;      // [kSortMaxDigits * tilesCap] digit-major tile histograms
    uint32_t* tilePrefix;  // [kSortMaxDigits * tilesCap] exclusive prefix over a segment's tiles per digit
    uint32_t* digitTotal;  // [segments * digits] per segment and digit
    uint32_t* tileStart;   // [numTiles + 1] first slot of each tile (tiles never straddle segments)
    uint32_t* tileSeg;     // [numTiles] segment of each tile
    uint32_t* segTile0;    // [segments + 1] first tile of each segment
    uint32_t numTiles;
    uint32_t tilesCap;
};

// Everything a kernel may need, passed by value.
// Launch variants: union records each neighbour pass stages in LDS per round.  They change LDS
// use (and so occupancy), never results.
struct Tuning {
    uint32_t pass1Round = 256;  // 128 or 256
    uint32_t pass2Round = 160;  // 128 or 160
    uint32_t driftRound = 160;  // 128, 160, 192 or 256
    uint32_t sortBits = 8;      // radix sort digit: 8 (8 passes) or 9 (7 passes)
    // wavefronts per block and per-wave stack entries of the gravity walk and of the neighbour
    // build: (2, 1024), (4, 1024), (4, 512) or (8, 512)
    uint32_t gravityWaves = 4;
    uint32_t gravityStack = 1024;
    uint32_t buildWaves = 4;
    uint32_t buildStack = 1024;
};

struct DeviceCtx {
    uint32_t slots;      // padded particle slots in use
    uint32_t blocks;     // slots / kBlock
    uint32_t stride64;   // element stride between the x, y, z parts of p64 / v64 (allocated slots)
    Segments seg;
    SegScalars ss;
    Persist cur;         // state in the current order
    Persist alt;         // permutation target
    Derived d;
    Tree t;
    Neighbours nb;
    Counters cnt;
    SortBuffers sb;
    Params prm;
    bool fp64;           // FP64 position/velocity accumulators
    bool strength;       // any material has strength enabled
    Tuning tune;
};

// The hold flag.  The neighbour build sets it when the table of a step came out incomplete (a
// bank, a union or the walk's stack too small).  From then on every kernel of the step sequence
// that would advance the state returns at once, and the permutation copies the state in place of
// gathering it, so the state, the tree and the step's dt stay those of the incomplete step while
// the graph or the queue of steps runs out.  The host then widens the table and finishes that
// step.  The flag only changes between launches on the context stream, so a kernel reads it once,
// uniformly, at its start.
__device__ __forceinline__ bool stepHeld(const DeviceCtx& c) {
    return __hip_atomic_load(&c.cnt.v[kCntHold], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0ull;
}

// sort.hip: stable LSD radix sort of (key, value) over `bits` key bits, then (segments > 1)
// one stable pass on the segment of each value.  Result in keyA/valA.
void sortPairs(const DeviceCtx& c, hipStream_t s);
// The context the Morton kernel writes its keys through: the buffers an odd number of sort
// passes starts from, so that the sorted keys and slots end in keyA / valA.
DeviceCtx sortInput(const DeviceCtx& c);

// tree.hip
void launchPrepareSort(const DeviceCtx& c, hipStream_t s);    // posNew = pos, bounds
void launchBoundsReduce(const DeviceCtx& c, hipStream_t s);  // per-segment bounds from block partials
void launchMorton(const DeviceCtx& c, hipStream_t s);
void launchPermuteKarras(const DeviceCtx& c, hipStream_t s);  // cur -> alt + topology
void launchMoments(const DeviceCtx& c, hipStream_t s);

// neighbours.hip
void launchNeighbours(const DeviceCtx& c, hipStream_t s);

// sph.hip
void launchPass1(const DeviceCtx& c, hipStream_t s);
void launchEos(const DeviceCtx& c, hipStream_t s);
void launchPass2(const DeviceCtx& c, hipStream_t s);

// gravity.hip
void launchGravity(const DeviceCtx& c, hipStream_t s);

// integrate.hip
void launchStepBegin(const DeviceCtx& c, hipStream_t s, float maxDt);
void launchStepEnd(const DeviceCtx& c, hipStream_t s);
void launchKick(const DeviceCtx& c, hipStream_t s, bool second, bool clampThermal);
void launchDrift(const DeviceCtx& c, hipStream_t s);
void launchStepBeginNoBounds(const DeviceCtx& c, hipStream_t s, float maxDt);
void launchDtCandidates(const DeviceCtx& c, hipStream_t s);   // dt_i min + radSum from the current state
void launchResetBounds(const DeviceCtx& c, hipStream_t s);
void launchInitIntegrator(const DeviceCtx& c, hipStream_t s); // radPrev = radSum, erad = time = dtPrev = 0
void launchClear(const DeviceCtx& c, hipStream_t s);
void launchEnergy(const DeviceCtx& c, hipStream_t s, double* dOut);

// telemetry.hip: largest-remnant classification.  out holds ASTRAEA_CLASS_STRIDE doubles per
// segment, seed one key per segment; rem, acc and body hold 8 doubles per segment.
void launchClassBodies(const DeviceCtx& c, hipStream_t s, double* out, unsigned long long* seed);
void launchRemnant(const DeviceCtx& c, hipStream_t s, const double* rem, double* acc);
void launchClassify(const DeviceCtx& c, hipStream_t s, const double* rem, const double* body, double* out);

// relax.hip: corrections of the damped relaxation.  sums and out hold 8 doubles per segment;
// anchor holds nAnchor temperatures (K) indexed by original id.
void launchRelaxCorrect(const DeviceCtx& c, hipStream_t s, double* sums, const float* anchor, uint32_t nAnchor,
                        double damping);
void launchReanchor(const DeviceCtx& c, hipStream_t s, const float* anchor, uint32_t nAnchor);
void launchAnchorFromEos(const DeviceCtx& c, hipStream_t s, float* anchor, uint32_t nAnchor);
void launchRelaxStats(const DeviceCtx& c, hipStream_t s, double* out);  // sum m, sum m v^2, max |x|^2

// render.hip: 20-byte particle records (5 dwords each) at outBase[segment] + local index, speed
// bytes; 40-byte tracer records (10 dwords each) at the tracer index.  com, livingCom and sysCom
// hold 3 floats per segment.
void launchPackRender(const DeviceCtx& c, hipStream_t s, const float* com, const uint32_t* outBase, uint32_t* records,
                      uint8_t* speed);
void launchExtractTracers(const DeviceCtx& c, hipStream_t s, const int32_t* tracerOfOid, uint32_t nOid,
                          const float4* rest, const uint8_t* ocean, uint32_t nTracers, const float* livingCom,
                          const float* sysCom, uint32_t* out);
void launchMarkTracers(const DeviceCtx& c, hipStream_t s, const int32_t* tracerOfOid, uint32_t nOid);

}  // namespace astraea_hip
