// sort.hip — stable LSD radix sort of (uint64 Morton key, uint32 slot) pairs for gfx950.
//
// Replaces cub::DeviceRadixSort::SortPairs over bits [0, 63) (engine/sph_gravity.cu:2028-2046).
// The result is exactly std::stable_sort by key: every pass is stable.
//
// One pass (8-bit digit) = three kernels:
//   hist    per 2048-element tile, digit counts (digit-major: counts[d * tiles + t]);
//   scan    one block per digit: exclusive prefix over tiles + digit totals;
//   scatter per tile: stable local ranks from wavefront ballots (each lane's digit is
//           matched against the wave with 8 ballots), per-wave running counts in LDS,
//           an LDS-staged reorder, then coalesced runs to HBM.
// Tiles are wave-contiguous: wave w of a tile owns kIPT consecutive 64-element rows, so the
// element order inside a tile is (wave, row, lane) and ranks stay stable.
// With several segments a final stable pass on the segment id groups the scenarios.
#include "launch.hpp"

namespace astraea_hip {
namespace {

constexpr int kTile = (int)kSortTile;
constexpr int kIPT = kTile / kBlock;
static_assert(kTile % kBlock == 0, "a tile is whole rows of the block");

__device__ __forceinline__ uint32_t digitOf(uint64_t key, uint32_t val, int shift, const uint32_t* slotSeg, bool seg) {
    return seg ? slotSeg[val] : (uint32_t)(key >> shift) & 0xffu;
}

// Lanes holding the same 8-bit digit: the match mask, from 8 ballots.
__device__ __forceinline__ uint64_t matchDigit(uint32_t d, uint64_t valid) {
    uint64_t m = valid;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t blockExclusiveScan(uint32_t v, uint32_t* scratch /* >= 4 */, uint32_t* total) {
    const int lane = laneId(), wave = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = (uint32_t)__shfl_up((int)inc, o, 64);
        if (lane >= o) inc += t;
    }
    if (lane == 63) scratch[wave] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kWavesPerBlock; ++w) {
        const uint32_t s = scratch[w];
        if (w < wave) pre += s;
        tot += s;
    }
    __syncthreads();
    if (total) *total = tot;
    return pre + inc - v;
}

__global__ __launch_bounds__(kBlock) void kHist(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                uint32_t n, int shift, const uint32_t* __restrict__ slotSeg, int segPass,
                                                uint32_t* __restrict__ counts, uint32_t tiles) {
    __shared__ uint32_t h[kWavesPerBlock][256];
    const int tid = threadIdx.x, wave = tid >> 6, lane = laneId();
    for (int k = tid; k < kWavesPerBlock * 256; k += kBlock) (&h[0][0])[k] = 0u;
    __syncthreads();
    const uint32_t tile = blockIdx.x;
    const uint32_t base = tile * kTile + wave * kWave * kIPT;
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t idx = base + k * kWave + lane;
        const bool v = idx < n;
        uint32_t d = 0;
        if (v) d = digitOf(keys[idx], vals[idx], shift, slotSeg, segPass != 0);
        const uint64_t m = matchDigit(d, ballot(v));
        const uint64_t lt = (1ull << lane) - 1ull;
        if (v && (m & lt) == 0ull) h[wave][d] += (uint32_t)__popcll(m);
    }
    __syncthreads();
    const uint32_t c = h[0][tid] + h[1][tid] + h[2][tid] + h[3][tid];
    counts[(size_t)tid * tiles + tile] = c;
}

// One block per digit: exclusive prefix of counts over tiles.
__global__ __launch_bounds__(kBlock) void kScan(const uint32_t* __restrict__ counts, uint32_t* __restrict__ prefix,
                                                uint32_t* __restrict__ totals, uint32_t tiles) {
    __shared__ uint32_t scratch[kWavesPerBlock];
    const uint32_t d = blockIdx.x;
    const uint32_t* src = counts + (size_t)d * tiles;
    uint32_t* dst = prefix + (size_t)d * tiles;
    uint32_t carry = 0;
    for (uint32_t b = 0; b < tiles; b += kBlock) {
        const uint32_t t = b + threadIdx.x;
        const uint32_t v = t < tiles ? src[t] : 0u;
        uint32_t tot;
        const uint32_t ex = blockExclusiveScan(v, scratch, &tot);
        if (t < tiles) dst[t] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) totals[d] = carry;
}

__global__ __launch_bounds__(kBlock) void kScatter(const uint64_t* __restrict__ keysIn, const uint32_t* __restrict__ valsIn,
                                                   uint64_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut,
                                                   uint32_t n, int shift, const uint32_t* __restrict__ slotSeg, int segPass,
                                                   const uint32_t* __restrict__ prefix,
                                                   const uint32_t* __restrict__ totals, uint32_t tiles) {
    __shared__ uint64_t sKey[kTile];
    __shared__ uint32_t sVal[kTile];
    __shared__ uint32_t sWave[kWavesPerBlock][256];
    __shared__ uint32_t sLocal[256];
    __shared__ uint32_t sGlobal[256];
    __shared__ uint32_t scratch[kWavesPerBlock];
    const int tid = threadIdx.x, wave = tid >> 6, lane = laneId();
    const bool seg = segPass != 0;
    const uint32_t tile = blockIdx.x;
    for (int k = tid; k < kWavesPerBlock * 256; k += kBlock) (&sWave[0][0])[k] = 0u;
    // global start of every digit for this tile
    {
        const uint32_t tot = totals[tid];
        const uint32_t start = blockExclusiveScan(tot, scratch, nullptr);
        sGlobal[tid] = start + prefix[(size_t)tid * tiles + tile];
    }
    __syncthreads();

    uint64_t key[kIPT];
    uint32_t val[kIPT], rank[kIPT], dig[kIPT];
    const uint32_t base = tile * kTile + wave * kWave * kIPT;
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t idx = base + k * kWave + lane;
        const bool v = idx < n;
        key[k] = v ? keysIn[idx] : 0ull;
        val[k] = v ? valsIn[idx] : 0u;
        dig[k] = v ? digitOf(key[k], val[k], shift, slotSeg, seg) : 256u;
    }
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const bool v = dig[k] < 256u;
        const uint64_t m = matchDigit(dig[k] & 0xffu, ballot(v));
        uint32_t r = 0;
        if (v) {
            const uint32_t before = sWave[wave][dig[k]];
            r = before + (uint32_t)__popcll(m & lt);
            if ((m & lt) == 0ull) sWave[wave][dig[k]] = before + (uint32_t)__popcll(m);
        }
        rank[k] = r;
    }
    __syncthreads();
    {
        // per digit: exclusive prefix over waves, tile total, local start
        uint32_t acc = 0;
#pragma unroll
        for (int w = 0; w < kWavesPerBlock; ++w) {
            const uint32_t c = sWave[w][tid];
            sWave[w][tid] = acc;
            acc += c;
        }
        sLocal[tid] = blockExclusiveScan(acc, scratch, nullptr);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        if (dig[k] < 256u) {
            const uint32_t p = sLocal[dig[k]] + sWave[wave][dig[k]] + rank[k];
            sKey[p] = key[k];
            sVal[p] = val[k];
        }
    }
    __syncthreads();
    const uint32_t inTile = min((uint32_t)kTile, n - tile * kTile);
#pragma unroll
    for (int e = 0; e < kIPT; ++e) {
        const uint32_t p = e * kBlock + tid;
        if (p < inTile) {
            const uint64_t kk = sKey[p];
            const uint32_t vv = sVal[p];
            const uint32_t d = digitOf(kk, vv, shift, slotSeg, seg);
            const uint32_t g = sGlobal[d] + (p - sLocal[d]);
            keysOut[g] = kk;
            valsOut[g] = vv;
        }
    }
}

}  // namespace

void sortPairs(const DeviceCtx& c, hipStream_t s) {
    const uint32_t n = c.slots;
    const uint32_t tiles = (n + kTile - 1) / kTile;
    uint64_t* kin = c.sb.keyA;
    uint64_t* kout = c.sb.keyB;
    uint32_t* vin = c.sb.valA;
    uint32_t* vout = c.sb.valB;
    const int passes = 8 + (c.seg.count > 1 ? 1 : 0);
    for (int p = 0; p < passes; ++p) {
        const int segPass = p == 8 ? 1 : 0;
        const int shift = segPass ? 0 : 8 * p;
        kHist<<<tiles, kBlock, 0, s>>>(kin, vin, n, shift, c.seg.slotSeg, segPass, c.sb.counts, tiles);
        kScan<<<256, kBlock, 0, s>>>(c.sb.counts, c.sb.tilePrefix, c.sb.digitTotal, tiles);
        kScatter<<<tiles, kBlock, 0, s>>>(kin, vin, kout, vout, n, shift, c.seg.slotSeg, segPass, c.sb.tilePrefix,
                                          c.sb.digitTotal, tiles);
        uint64_t* tk = kin; kin = kout; kout = tk;
        uint32_t* tv = vin; vin = vout; vout = tv;
    }
    // an even number of passes leaves the result in keyA/valA; after an odd number copy it back
    if (passes & 1) {
        (void)hipMemcpyAsync(c.sb.keyA, kin, sizeof(uint64_t) * n, hipMemcpyDeviceToDevice, s);
        (void)hipMemcpyAsync(c.sb.valA, vin, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, s);
    }
}

}  // namespace astraea_hip
