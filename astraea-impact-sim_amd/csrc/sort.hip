// sort.hip — stable LSD radix sort of (uint64 Morton key, uint32 slot) pairs for gfx950.
//
// Replaces cub::DeviceRadixSort::SortPairs over bits [0, 63) (engine/sph_gravity.cu:2028-2046).
// The result is exactly std::stable_sort by key within every segment: every pass is stable.
//
// Each segment (scenario) is sorted inside its own slot range, which its particles never leave:
// the tiles are cut along segment boundaries (a segment's last tile may be short), a tile's
// elements go to its own segment's range, and no pass has to group the segments afterwards.
// One pass (a BITS-bit digit, 8 or 9: 8 or 7 passes over the 63 key bits) = three kernels:
//   hist    per tile of up to 2048 elements, digit counts (digit-major: counts[d * tilesCap + t]);
//   scan    one wavefront per (digit, segment): exclusive prefix over the segment's tiles + the
//           segment's digit total;
//   scatter per tile: stable local ranks from wavefront ballots (each lane's digit is
//           matched against the wave with BITS ballots), per-wave running counts in LDS,
//           an LDS-staged reorder, then coalesced runs to HBM.
// Tiles are wave-contiguous: wave w of a tile owns kIPT consecutive 64-element rows, so the
// element order inside a tile is (wave, row, lane) and ranks stay stable.  The passes alternate
// between the two buffers and end in keyA/valA: with an odd number of passes the Morton kernel
// writes its keys to keyB/valB (sortInput).
#include "launch.hpp"

namespace astraea_hip {
namespace {

constexpr int kTile = (int)kSortTile;
constexpr int kIPT = kTile / kBlock;
static_assert(kTile % kBlock == 0, "a tile is whole rows of the block");

// Lanes holding the same BITS-bit digit: the match mask, from BITS ballots.
template <int BITS>
__device__ __forceinline__ uint64_t matchDigit(uint32_t d, uint64_t valid) {
    uint64_t m = valid;
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
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

// Exclusive scan over the DIG digits of `in` (LDS or global), thread t owning digits
// t * DIG / 256 .. (t + 1) * DIG / 256 - 1; writes out[d] (LDS).
template <int DIG>
__device__ __forceinline__ void digitExclusiveScan(const uint32_t* in, uint32_t* out, uint32_t* scratch) {
    constexpr int PER = DIG / kBlock;
    static_assert(DIG % kBlock == 0, "digits are whole multiples of the block");
    const int tid = threadIdx.x;
    uint32_t v[PER], sum = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        v[k] = in[tid * PER + k];
        sum += v[k];
    }
    uint32_t run = blockExclusiveScan(sum, scratch, nullptr);
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        out[tid * PER + k] = run;
        run += v[k];
    }
}

template <int BITS>
__global__ __launch_bounds__(kBlock) void kHist(const uint64_t* __restrict__ keys, uint32_t* __restrict__ counts,
                                                const uint32_t* __restrict__ tileStart, uint32_t tilesCap, int shift) {
    constexpr int DIG = 1 << BITS;
    __shared__ uint32_t h[kWavesPerBlock][DIG];
    const int tid = threadIdx.x, wave = tid >> 6, lane = laneId();
    for (int k = tid; k < kWavesPerBlock * DIG; k += kBlock) (&h[0][0])[k] = 0u;
    __syncthreads();
    const uint32_t tile = xcdBlock(blockIdx.x, gridDim.x);  // the scatter's tile order
    const uint32_t end = tileStart[tile + 1];
    const uint32_t base = tileStart[tile] + wave * kWave * kIPT;
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t idx = base + k * kWave + lane;
        const bool v = idx < end;
        // counts only (no ranks): an LDS atomic per element, where matching the digit across the
        // wavefront would take BITS ballots
        if (v) atomicAdd(&h[wave][(uint32_t)(keys[idx] >> shift) & (uint32_t)(DIG - 1)], 1u);
    }
    __syncthreads();
    for (int d = tid; d < DIG; d += kBlock) {
        uint32_t c = 0;
#pragma unroll
        for (int w = 0; w < kWavesPerBlock; ++w) c += h[w][d];
        counts[(size_t)d * tilesCap + tile] = c;
    }
}

// One wavefront per (digit, segment): the exclusive prefix of the digit's counts over the
// segment's tiles, 64 at a time, and the segment's total of the digit.
__global__ __launch_bounds__(kBlock) void kScan(const uint32_t* __restrict__ counts, uint32_t* __restrict__ prefix,
                                                uint32_t* __restrict__ totals, const uint32_t* __restrict__ segTile0,
                                                uint32_t segments, uint32_t tilesCap, uint32_t digits) {
    const int lane = laneId();
    const uint32_t pair = uniformU32(blockIdx.x * (uint32_t)kWavesPerBlock + (threadIdx.x >> 6));
    if (pair >= digits * segments) return;
    const uint32_t d = pair / segments, s = pair % segments;
    const uint32_t* src = counts + (size_t)d * tilesCap;
    uint32_t* dst = prefix + (size_t)d * tilesCap;
    const uint32_t t0 = uniformU32(segTile0[s]), t1 = uniformU32(segTile0[s + 1]);
    uint32_t carry = 0;
    for (uint32_t b = t0; b < t1; b += kWave) {
        const uint32_t t = b + (uint32_t)lane;
        const uint32_t v = t < t1 ? src[t] : 0u;
        uint32_t inc = v;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t u = (uint32_t)__shfl_up((int)inc, o, 64);
            if (lane >= o) inc += u;
        }
        if (t < t1) dst[t] = carry + inc - v;
        carry += (uint32_t)__builtin_amdgcn_readlane((int)inc, kWave - 1);
    }
    if (lane == 0) totals[(size_t)s * digits + d] = carry;
}

template <int BITS>
__global__ __launch_bounds__(kBlock) void kScatter(const uint64_t* __restrict__ keysIn, const uint32_t* __restrict__ valsIn,
                                                   uint64_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut,
                                                   const uint32_t* __restrict__ tileStart,
                                                   const uint32_t* __restrict__ tileSeg,
                                                   const uint32_t* __restrict__ segOff, const uint32_t* __restrict__ prefix,
                                                   const uint32_t* __restrict__ totals, uint32_t tilesCap, int shift) {
    constexpr int DIG = 1 << BITS;
    // the per-wave digit counts are dead once every element's destination is in registers, so
    // they share their LDS with the staged keys
    __shared__ union {
        uint64_t key[kTile];
        uint32_t wave[kWavesPerBlock][DIG];
    } sKW;
    static_assert(sizeof(uint32_t) * kWavesPerBlock * DIG <= sizeof(uint64_t) * kTile, "the counts fit the keys' space");
    uint64_t* const sKey = sKW.key;
    uint32_t (*const sWave)[DIG] = sKW.wave;
    __shared__ uint32_t sVal[kTile];
    __shared__ uint32_t sLocal[DIG];
    __shared__ uint32_t sGlobal[DIG];
    __shared__ uint32_t scratch[kWavesPerBlock];
    const int tid = threadIdx.x, wave = tid >> 6, lane = laneId();
    // consecutive tiles on one XCD: tile t's run of a digit ends where tile t + 1's begins, and the
    // cache line they share is completed in one L2 instead of reaching HBM as two partial writes
    const uint32_t tile = xcdBlock(blockIdx.x, gridDim.x);
    const uint32_t seg = tileSeg[tile];
    const uint32_t t0 = tileStart[tile], end = tileStart[tile + 1];
    for (int k = tid; k < kWavesPerBlock * DIG; k += kBlock) (&sWave[0][0])[k] = 0u;
    // start of every digit for this tile: the segment's offset, the digits below in the segment,
    // the same digit in the segment's earlier tiles
    digitExclusiveScan<DIG>(totals + (size_t)seg * DIG, sGlobal, scratch);
    __syncthreads();
    {
        const uint32_t so = segOff[seg];
        for (int d = tid; d < DIG; d += kBlock) sGlobal[d] += so + prefix[(size_t)d * tilesCap + tile];
    }
    __syncthreads();

    uint64_t key[kIPT];
    uint32_t val[kIPT], rank[kIPT], dig[kIPT];
    const uint32_t base = t0 + wave * kWave * kIPT;
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const uint32_t idx = base + k * kWave + lane;
        const bool v = idx < end;
        key[k] = v ? keysIn[idx] : 0ull;
        val[k] = v ? valsIn[idx] : 0u;
        dig[k] = v ? (uint32_t)(key[k] >> shift) & (uint32_t)(DIG - 1) : (uint32_t)DIG;
    }
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        const bool v = dig[k] < (uint32_t)DIG;
        const uint64_t m = matchDigit<BITS>(dig[k] & (uint32_t)(DIG - 1), ballot(v));
        uint32_t r = 0;
        if (v) {
            const uint32_t before = sWave[wave][dig[k]];
            r = before + (uint32_t)__popcll(m & lt);
            if ((m & lt) == 0ull) sWave[wave][dig[k]] = before + (uint32_t)__popcll(m);
        }
        rank[k] = r;
    }
    __syncthreads();
    // per digit: exclusive prefix over waves (in place) and the tile's count (sLocal, then its
    // exclusive prefix over digits: the digit's start inside the tile)
    for (int d = tid; d < DIG; d += kBlock) {
        uint32_t acc = 0;
#pragma unroll
        for (int w = 0; w < kWavesPerBlock; ++w) {
            const uint32_t c = sWave[w][d];
            sWave[w][d] = acc;
            acc += c;
        }
        sLocal[d] = acc;
    }
    __syncthreads();
    digitExclusiveScan<DIG>(sLocal, sLocal, scratch);
    __syncthreads();
    uint32_t pos[kIPT];
#pragma unroll
    for (int k = 0; k < kIPT; ++k)
        pos[k] = dig[k] < (uint32_t)DIG ? sLocal[dig[k]] + sWave[wave][dig[k]] + rank[k] : 0u;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kIPT; ++k) {
        if (dig[k] < (uint32_t)DIG) {
            sKey[pos[k]] = key[k];
            sVal[pos[k]] = val[k];
        }
    }
    __syncthreads();
    const uint32_t inTile = end - t0;
#pragma unroll
    for (int e = 0; e < kIPT; ++e) {
        const uint32_t p = e * kBlock + tid;
        if (p < inTile) {
            const uint64_t kk = sKey[p];
            const uint32_t vv = sVal[p];
            const uint32_t d = (uint32_t)(kk >> shift) & (uint32_t)(DIG - 1);
            const uint32_t g = sGlobal[d] + (p - sLocal[d]);
            keysOut[g] = kk;
            valsOut[g] = vv;
        }
    }
}

template <int BITS>
void sortPasses(const DeviceCtx& c, hipStream_t s) {
    constexpr int DIG = 1 << BITS;
    constexpr int passes = (63 + BITS - 1) / BITS;
    const uint32_t tiles = c.sb.numTiles;
    uint64_t* kin = passes % 2 ? c.sb.keyB : c.sb.keyA;
    uint64_t* kout = passes % 2 ? c.sb.keyA : c.sb.keyB;
    uint32_t* vin = passes % 2 ? c.sb.valB : c.sb.valA;
    uint32_t* vout = passes % 2 ? c.sb.valA : c.sb.valB;
    for (int p = 0; p < passes; ++p) {
        const int shift = BITS * p;
        kHist<BITS><<<tiles, kBlock, 0, s>>>(kin, c.sb.counts, c.sb.tileStart, c.sb.tilesCap, shift);
        kScan<<<(DIG * c.seg.count + kWavesPerBlock - 1) / kWavesPerBlock, kBlock, 0, s>>>(
            c.sb.counts, c.sb.tilePrefix, c.sb.digitTotal, c.sb.segTile0, c.seg.count, c.sb.tilesCap, (uint32_t)DIG);
        kScatter<BITS><<<tiles, kBlock, 0, s>>>(kin, vin, kout, vout, c.sb.tileStart, c.sb.tileSeg, c.seg.off,
                                                c.sb.tilePrefix, c.sb.digitTotal, c.sb.tilesCap, shift);
        uint64_t* tk = kin; kin = kout; kout = tk;
        uint32_t* tv = vin; vin = vout; vout = tv;
    }
}

}  // namespace

DeviceCtx sortInput(const DeviceCtx& c) {
    DeviceCtx m = c;
    if (sortPassesOf(c.tune.sortBits) % 2) {
        m.sb.keyA = c.sb.keyB;
        m.sb.valA = c.sb.valB;
    }
    return m;
}

void sortPairs(const DeviceCtx& c, hipStream_t s) {
    if (c.tune.sortBits == 9u) sortPasses<9>(c, s);
    else sortPasses<8>(c, s);
}

}  // namespace astraea_hip
