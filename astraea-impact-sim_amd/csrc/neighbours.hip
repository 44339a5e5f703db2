// neighbours.hip — per-particle neighbour table built once per step.
//
// The reference walks the tree three times per step, once per neighbour pass (pass 1
// sph_gravity.cu:645-788, pass 2 971-1287, XSPH drift 1376-1489), and evaluates the pair
// physics for a whole warp whenever any lane is in range.  Here one table is built right
// after the tree, holding for every particle i all j != i with
//     |x_i - x_j|^2 < max(Hb_i, Hb_j)^2,   Hb = supportBound(h, h0) >= every support H' that
// pass 1's h update can produce.  It is a superset of each pass's set:
//     pass 1   r < H_i             (H of the pre-update h; H <= Hb)
//     pass 2   r < max(H_i', H_j')  (post-update h)
//     drift    r < max(H_i', H_j')  of the next step (same positions, same order)
// and every pass re-applies its own test with the same squared-distance expression, so the
// sets are exactly the reference's.  Each lane then loops over its own list: the pair
// physics runs only for real pairs.
//
// Layout (WaveTable): each wavefront writes the union of its particles' neighbour sets (the hit
// candidates, in the order the walk found them, dealt out to kBanks banks) and, for every
// particle and bank, the ascending 16-bit positions of its neighbours in that bank, in 16-byte
// chunks of eight written whole.  A wavefront first builds into the regular region,
// `stride` entries per slot and `unionCap` union ids.  One that needs more (a particle with many
// neighbours, as an envelope particle of a gas giant with h ten times its core's) is listed, and a
// second kernel builds it again into the spill pool with exactly the width it needs, so memory
// follows the entries the particles have, not the worst particle times every slot.  Only a pool
// too small for the listed wavefronts (or a union beyond 65536 ids, or the walk's stack) leaves
// the table incomplete: then the hold flag (stepHeld) stops the rest of the step sequence, and the
// host grows the pool, builds the table again from the same tree and finishes the step.
#include "launch.hpp"

namespace astraea_hip {
namespace {

constexpr int kCCap = 256;                   // per-wave candidate records
constexpr int kFlushAt = kCCap - 2 * kWave;  // one round lists at most 128 leaves
constexpr int kBucket = 16;                  // an open node with this many leaves or fewer is listed whole
constexpr int kBCap = 192;                   // per-wave listed buckets (a round adds at most 128)
constexpr int kBucketFlushAt = kBCap - 2 * kWave;
constexpr int kSlotBits = 28;                // a bucket: first leaf slot | (leaves - 1) << kSlotBits
constexpr uint32_t kBucketSlotMask = (1u << kSlotBits) - 1u;
static_assert(kBucket <= (1 << (32 - kSlotBits)) && kWave % kBucket == 0, "a bucket's leaf count fits the top bits");

template <int SCAP>  // per-wave node stack: 1024 or 512 entries (a launch variant)
struct WaveLds {
    int stack[SCAP];
    float4 cand[kCCap];  // candidate particle (x, y, z, squared support bound)
    int candId[kCCap];
    uint32_t bucket[kBCap];
    int selfAt[kWave];   // where lane l's own leaf sits among the listed candidates (-1: not listed)
};

// Exclusive prefix parity of a 32-bit mask: bit u is the parity of the set bits of x below u.
__device__ __forceinline__ uint32_t prefixParityBelow(uint32_t x) {
    uint32_t p = x;
    p ^= p << 1;
    p ^= p << 2;
    p ^= p << 4;
    p ^= p << 8;
    p ^= p << 16;
    return p ^ x;
}

// Four 16-bit entries' worth of a lane's current chunk of one bank, filled by shifting: after
// eight appends entry 0 is the low half of `a`, entry 7 the high half of `d` (the table's order).
struct ChunkAcc {
    uint32_t a, b, c, d;
};
__device__ __forceinline__ void appendEntry(ChunkAcc& k, uint32_t pos) {
    k.a = __builtin_amdgcn_alignbit(k.b, k.a, 16u);
    k.b = __builtin_amdgcn_alignbit(k.c, k.b, 16u);
    k.c = __builtin_amdgcn_alignbit(k.d, k.c, 16u);
    k.d = __builtin_amdgcn_alignbit(pos, k.d, 16u);
}

// Wide walk: the wave expands up to 64 stack nodes per round, one per lane.  A node must be opened
// when some group particle's sphere of radius max(Hb_p, node's max Hb) meets its box.  The group is
// cut into four sub-groups of 16 Morton-consecutive particles.  In FP32 as in exact arithmetic
// (rounding is monotone) the box distance between a sub-group's box and the node's box is a lower
// bound of the box distance of each of its particles, so a node out of reach of every sub-group
// is closed for every particle.  A compact sub-group (box sides within twice its largest support)
// opens the nodes within its reach directly: the leaves are tested pair by pair afterwards, so the
// table stays exact.  A sub-group that straddles a jump of the Morton curve has a box far larger
// than its particles' spheres; nodes that only such a sub-group reaches are tested exactly, one
// node at a time across the lanes (lane = particle), with the node's box broadcast from the lane
// that loaded it.  Leaf children are loaded by their lanes in parallel, kept when within reach of
// a sub-group's box (box distance is a lower bound of every member's distance, in FP32 as well,
// so no hit is lost), and staged in LDS, where every lane tests its own particle against them:
// 32 candidates into a hit mask, then the lane appends its hits in candidate order, so the
// bookkeeping of an entry runs once per hit rather than once per candidate.  The candidates hit by
// any lane form the wave's union list; a lane's entry is the hit's position in it.
//
// Entries are collected per lane and bank in registers (ChunkAcc) and each chunk of eight is
// written once, as one 16-byte store, when it fills (a partial last chunk when the walk ends).
// Which bank a hit goes to follows from its rank among the 32 candidates' union members (rank
// mod 4, from two prefix parities of the wave-uniform union mask), so the banks are handled one at
// a time with their registers known at compile time.
struct WaveBuild {
    uint64_t cnt4;    // entries per bank of this lane's particle: bank b in bits 16 b .. 16 b + 15
    uint32_t ulen;    // length of the wavefront's union (uncapped)
    bool stackFull;
};

// Builds the table of wavefront gw (slots 64 gw .. 64 gw + 63) into `list` (64 lanes x K entries in
// chunks, WaveTable) and `U` (UC union ids).  Every lane of the wavefront calls it.
template <int SCAP>
__device__ __forceinline__ WaveBuild buildWave(const DeviceCtx& c, WaveLds<SCAP>& L, uint32_t gw, int lane, uint16_t* list,
                                               uint32_t K, uint32_t* U, uint32_t UC) {
    const uint32_t i = gw * kWave + (uint32_t)lane;
    const uint32_t sg = uniformU32(c.seg.slotSeg[gw * kWave]);
    const uint32_t off = uniformU32(c.seg.off[sg]), n = uniformU32(c.seg.n[sg]);
    const bool active = i - off < n;
    const uint32_t Kb = K / kBanks, perBank = K / (8u * kBanks);
    uint4* out = reinterpret_cast<uint4*>(list) + lane;
    uint64_t cnt4 = 0;
    uint32_t ulen = 0;
    bool stackFull = false;
    const uint32_t UCb = UC / kBanks;
    const uint64_t actMask = ballot(active);
    if (actMask != 0ull && n >= 2u) {
        float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float Hb = 0.0f;
        if (active) {
            pi = c.cur.pos[i];
            const float4 hu = c.cur.hu[i];
            Hb = supportBound(hu.x, hu.y);
        }
        const float Hb2 = Hb * Hb;
        // sub-group boxes (lanes 16q .. 16q+15); a sub-group without particles gets an empty box
        float sl[4][3], sh[4][3], sH[4];
        bool compact[4];
        {
            float v[7] = {active ? pi.x : 3.0e38f, active ? pi.y : 3.0e38f, active ? pi.z : 3.0e38f,
                          active ? pi.x : -3.0e38f, active ? pi.y : -3.0e38f, active ? pi.z : -3.0e38f,
                          active ? Hb : 0.0f};
            for (int o = 1; o < 16; o <<= 1) {
                for (int a = 0; a < 3; ++a) v[a] = fminf(v[a], __shfl_xor(v[a], o, 64));
                for (int a = 3; a < 7; ++a) v[a] = fmaxf(v[a], __shfl_xor(v[a], o, 64));
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    sl[q][a] = readLaneF(v[a], 16 * q);
                    sh[q][a] = readLaneF(v[3 + a], 16 * q);
                }
                sH[q] = readLaneF(v[6], 16 * q);
                const float side = fmaxf(fmaxf(sh[q][0] - sl[q][0], sh[q][1] - sl[q][1]), sh[q][2] - sl[q][2]);
                compact[q] = side <= 2.0f * sH[q];  // an empty sub-group is compact and reaches nothing
            }
        }
        const uint64_t lt = (1ull << lane) - 1ull;
        int sp = 1, nl = 0, nr = 0;
        ChunkAcc acc[kBanks] = {};
        L.stack[0] = (int)off;
        L.selfAt[lane] = -1;
        __builtin_amdgcn_wave_barrier();
        // a leaf j as a candidate record, kept when within reach of a sub-group's box (an empty
        // sub-group's inverted box is out of reach of everything)
        const auto leafCandidate = [&](uint32_t j, float4& q) -> bool {
            const float4 pj = c.cur.pos[j];
            const float4 hj = c.cur.hu[j];
            const float hbj = supportBound(hj.x, hj.y);
            q = make_float4(pj.x, pj.y, pj.z, hbj * hbj);  // max(a, b)^2 = max(a^2, b^2) in FP32 too
            bool keep = false;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const float reach = fmaxf(sH[g], hbj);
                const float gx = fmaxf(0.0f, fmaxf(sl[g][0] - q.x, q.x - sh[g][0]));
                const float gy = fmaxf(0.0f, fmaxf(sl[g][1] - q.y, q.y - sh[g][1]));
                const float gz = fmaxf(0.0f, fmaxf(sl[g][2] - q.z, q.z - sh[g][2]));
                keep = keep || dist2f(gx, gy, gz) < reach * reach;
            }
            return keep;
        };
        const auto listLeaves = [&](bool keep, const float4& q, uint32_t j) {
            const uint64_t mList = ballot(keep);
            if (keep) {
                const int at = nl + __popcll(mList & lt);
                L.cand[at] = q;
                L.candId[at] = (int)j;
                const uint32_t rel = j - gw * kWave;  // the wavefront's own particle: lane rel
                if (rel < (uint32_t)kWave) L.selfAt[rel] = at;
            }
            nl += __popcll(mList);
        };
        // candidates: every lane tests its own particle (r^2 < max(Hb_i, Hb_j)^2), 32 at a time;
        // slots past nl hold stale records and are masked off; the candidates hit by any
        // lane (a wave-uniform mask from the ballots) join the union, in candidate order
        const auto testCandidates = [&]() {
            __builtin_amdgcn_wave_barrier();
            const int selfAt = L.selfAt[lane];  // a leaf is listed once per walk
            for (int c0 = 0; c0 < nl; c0 += 32) {
                uint32_t m = 0u, any = 0u;
                // eight tests per unrolled body: fully unrolled, the 32 compare masks held at once
                // pushed the sub-group boxes out of the scalar registers (50 spilled, reloaded in
                // every walk round)
#pragma unroll 8
                for (int u = 0; u < 32; ++u) {
                    const float4 q = L.cand[c0 + u];
                    const float r2 = dist2f(pi.x - q.x, pi.y - q.y, pi.z - q.z);
                    const bool hit = r2 < maxPos(Hb2, q.w);  // both are squares
                    m |= hit ? (1u << u) : 0u;
                    any |= (ballot(hit) & actMask) != 0ull ? (1u << u) : 0u;
                }
                const int left = nl - c0;
                if (left < 32) {
                    m &= (1u << left) - 1u;
                    any &= (1u << left) - 1u;
                }
                if (!active) m = 0u;
                if (selfAt >= c0 && selfAt < c0 + 32) m &= ~(1u << (selfAt - c0));  // j != i
                if (lane < 32 && ((any >> lane) & 1u)) {
                    const uint32_t q = ulen + (uint32_t)__popc(any & ((1u << lane) - 1u));
                    if (q < UC) U[(q % kBanks) * UCb + q / kBanks] = (uint32_t)L.candId[c0 + lane];
                }
                // union member u has rank r(u) among the members; it goes to bank (ulen + r) % 4 at
                // position (ulen + r) / 4.  Label r % 4 = 2 (parity of the odd-ranked members below
                // u) + (parity of the members below u).
                const uint32_t par1 = prefixParityBelow(any);
                const uint32_t par2 = prefixParityBelow(any & par1);
                const uint32_t lab[4] = {any & ~par1 & ~par2, any & par1 & ~par2, any & ~par1 & par2, any & par1 & par2};
#pragma unroll
                for (uint32_t bk = 0; bk < kBanks; ++bk) {
                    const uint32_t l = (bk - ulen) & 3u;
                    const uint32_t mb = lab[l];
                    const uint32_t first = (ulen + l) / kBanks;  // position of mb's first member
                    uint32_t hits = m & mb;
                    const uint32_t sh = bk << 4;
                    uint32_t cb = (uint32_t)(cnt4 >> sh) & 0xffffu;
                    cnt4 += (uint64_t)__popc(hits) << sh;
                    while (hits != 0u) {
                        const int u = __builtin_ctz(hits);
                        hits &= hits - 1u;
                        appendEntry(acc[bk], first + (uint32_t)__popc(mb & ((1u << u) - 1u)));
                        ++cb;
                        // a full chunk inside the bank's capacity is written; past the capacity (or
                        // past the union's) the entries are only counted: the wavefront is built
                        // again in the spill pool with the width the counts ask for
                        if ((cb & 7u) == 0u && cb <= Kb)
                            out[64ull * (bk * perBank + (cb >> 3) - 1u)] =
                                make_uint4(acc[bk].a, acc[bk].b, acc[bk].c, acc[bk].d);
                    }
                }
                ulen += (uint32_t)__popc(any);
            }
            L.selfAt[lane] = -1;
            nl = 0;
            __builtin_amdgcn_wave_barrier();
        };
        for (;;) {
            while (sp > 0 && nl <= kFlushAt && nr <= kBucketFlushAt) {
                const int k = sp < kWave ? sp : kWave;
                sp -= k;
                // up to 32 nodes: a lane per child (lanes 2t, 2t + 1 take node t's two children) in
                // one pass; more: a lane per node, its two children in two passes
                const bool paired = k <= kWave / 2;
                const int at = paired ? lane >> 1 : lane;
                const bool has = at < k;
                int2 ch = make_int2(0, 0);
                if (has) ch = c.t.child[L.stack[sp + at]];
                __builtin_amdgcn_wave_barrier();
                const int passes = paired ? 1 : 2;
                for (int side = 0; side < passes; ++side) {
                    const int cc = (paired ? (lane & 1) : side) == 0 ? ch.x : ch.y;
                    const bool isInt = has && cc >= 0;
                    bool isLeaf = has && cc < 0;
                    float4 q = make_float4(0.0f, 0.0f, 0.0f, 0.0f), mn = q, mx = q;
                    bool open = false, undecided = false;
                    int far = 0;
                    if (isInt) {
                        mn = c.t.bmin[cc];
                        mx = c.t.bmax[cc];
                        far = c.t.other[cc];
#pragma unroll
                        for (int g = 0; g < 4; ++g) {
                            const float reach = fmaxf(sH[g], mx.w);
                            const float gx = fmaxf(0.0f, fmaxf(mn.x - sh[g][0], sl[g][0] - mx.x));
                            const float gy = fmaxf(0.0f, fmaxf(mn.y - sh[g][1], sl[g][1] - mx.y));
                            const float gz = fmaxf(0.0f, fmaxf(mn.z - sh[g][2], sl[g][2] - mx.z));
                            const bool hit = dist2f(gx, gy, gz) <= reach * reach;
                            if (compact[g]) open = open || hit;
                            else undecided = undecided || hit;
                        }
                        undecided = undecided && !open;
                    } else if (isLeaf) {
                        isLeaf = leafCandidate((uint32_t)~cc, q);
                    }
                    // exact test of the nodes only straddling sub-groups reach: lane = particle
                    uint64_t um = ballot(undecided);
                    while (um != 0ull) {
                        const int src = (int)__builtin_ctzll(um);
                        um &= um - 1ull;
                        const float4 bmn = make_float4(readLaneF(mn.x, src), readLaneF(mn.y, src), readLaneF(mn.z, src), 0.0f);
                        const float4 bmx = make_float4(readLaneF(mx.x, src), readLaneF(mx.y, src), readLaneF(mx.z, src), 0.0f);
                        const float rr = fmaxf(Hb, readLaneF(mx.w, src));
                        const bool hit = active && boxDist2f(pi.x, pi.y, pi.z, bmn, bmx) <= rr * rr;
                        const bool any = ballot(hit) != 0ull;
                        if (lane == src) open = any;
                    }
                    // an open node of at most kBucket leaves is listed whole instead of walked
                    const int span = far > cc ? far - cc : cc - far;  // leaves - 1
                    const bool bucket = open && span < kBucket;
                    open = open && !bucket;
                    const uint64_t mBucket = ballot(bucket);
                    if (bucket) L.bucket[nr + __popcll(mBucket & lt)] = (uint32_t)min(cc, far) | ((uint32_t)span << kSlotBits);
                    nr += __popcll(mBucket);
                    const uint64_t mPush = ballot(open);
                    const int pushAt = sp + __popcll(mPush & lt);
                    if (open && pushAt >= SCAP) {
                        open = false;
                        stackFull = true;
                    }
                    if (open) L.stack[pushAt] = cc;
                    const int nPush = __popcll(mPush);
                    sp = sp + nPush < SCAP ? sp + nPush : SCAP;
                    listLeaves(isLeaf, q, (uint32_t)~cc);
                }
                __builtin_amdgcn_wave_barrier();
            }
            // the listed buckets, kWave / kBucket at a time: lane kBucket r + e takes leaf e of bucket r
            for (int r0 = 0; r0 < nr; r0 += kWave / kBucket) {
                if (nl > kCCap - kWave) testCandidates();
                const int r = r0 + lane / kBucket, e = lane % kBucket;
                const uint32_t rec = r < nr ? L.bucket[r] : 0u;
                const uint32_t j = (rec & kBucketSlotMask) + (uint32_t)e;
                float4 q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                bool keep = r < nr && (uint32_t)e <= (rec >> kSlotBits);
                if (keep) keep = leafCandidate(j, q);
                listLeaves(keep, q, j);
            }
            nr = 0;
            if (nl > kFlushAt || sp == 0) testCandidates();
            if (sp == 0) break;
        }
        // the partial last chunk of every bank, its entries shifted down to half-words 0, 1, ...
        // (chunk cb / 8 lies inside the bank's capacity when cb < Kb)
#pragma unroll
        for (uint32_t bk = 0; bk < kBanks; ++bk) {
            const uint32_t cb = (uint32_t)(cnt4 >> (bk << 4)) & 0xffffu;
            if ((cb & 7u) != 0u && cb < Kb) {
                for (uint32_t t = cb & 7u; t < 8u; ++t) appendEntry(acc[bk], 0u);
                out[64ull * (bk * perBank + (cb >> 3))] = make_uint4(acc[bk].a, acc[bk].b, acc[bk].c, acc[bk].d);
            }
        }
    }
    return WaveBuild{cnt4, ulen, stackFull};
}

// Per-slot counts (capped at the bank's capacity) and per-wavefront sizes of a build; true when
// every bank and the union fit.
__device__ __forceinline__ bool finishWave(const DeviceCtx& c, uint32_t gw, int lane, const WaveBuild& w, uint32_t K,
                                           uint32_t UC) {
    const uint32_t i = gw * kWave + (uint32_t)lane;
    const uint32_t Kb = K / kBanks;
    uint32_t cnt = 0, maxBank = 0, stored[kBanks];
#pragma unroll
    for (uint32_t bk = 0; bk < kBanks; ++bk) {
        const uint32_t cb = (uint32_t)(w.cnt4 >> (16u * bk)) & 0xffffu;
        cnt += cb;
        maxBank = max(maxBank, cb);
        stored[bk] = min(cb, Kb);
    }
    c.nb.count[i] = make_uint2(stored[0] | (stored[1] << 16), stored[2] | (stored[3] << 16));
    uint32_t sum = cnt;
    for (int o = 32; o > 0; o >>= 1) sum += (uint32_t)__shfl_xor((int)sum, o, 64);
    const uint32_t mxc = waveMaxU(cnt);
    const uint32_t want = waveMaxU(maxBank) * kBanks;  // the width the fullest bank needs
    const bool full = ballot(w.stackFull) != 0ull;
    if (lane == 0) {
        c.nb.unionLen[gw] = w.ulen;
        c.nb.waveSum[gw] = sum;
        c.nb.waveMax[gw] = mxc;
        c.nb.waveNeed[gw] = (want + 31u) & ~31u;
        atomicMax(&c.cnt.v[7], (unsigned long long)want);
        atomicMax(&c.cnt.v[5], (unsigned long long)w.ulen);
        if (full) {
            atomicOr(&c.cnt.v[6], 1ull);
            atomicOr(&c.cnt.v[kCntHold], 1ull);
        }
        if (w.ulen > (1u << 16)) {  // beyond what 16-bit bank positions address
            atomicOr(&c.cnt.v[3], 1ull);
            atomicOr(&c.cnt.v[kCntHold], 1ull);
        }
    }
    return want <= K && w.ulen <= UC;
}

// First build: every wavefront into the regular region; the ones that need more are listed.  A
// block is W wavefronts with SCAP-entry stacks (launch variants: how many neighbouring walks share a
// CU's L1 and a block's LDS; a full stack makes the build incomplete, and the host builds again).
template <int W, int SCAP>
__global__ __launch_bounds__(W * kWave) void kNeighbours(DeviceCtx c) {
    __shared__ WaveLds<SCAP> lds[W];
    if (stepHeld(c)) return;
    const uint32_t b = xcdBlock(blockIdx.x, gridDim.x);
    const int wave = threadIdx.x >> 6, lane = laneId();
    const uint32_t gw = b * (uint32_t)W + (uint32_t)wave;
    if (gw * (uint32_t)kWave >= c.slots) return;  // the last block of a grid over a multiple of 256 slots
    const uint32_t K = c.nb.stride, UC = c.nb.unionCap;
    uint16_t* list = c.nb.list + (uint64_t)gw * (64ull * K);
    uint32_t* U = c.nb.unionIds + (uint64_t)gw * UC;
    const WaveBuild w = buildWave(c, lds[wave], gw, lane, list, K, U, UC);
    const bool fits = finishWave(c, gw, lane, w, K, UC);
    if (lane == 0) {
        if (fits) {
            c.nb.tab[gw] = WaveTable{list, U, K, UC};
        } else {
            const unsigned long long at = atomicAdd(&c.nb.q[0], 1ull);
            c.nb.spill[at] = gw;
        }
    }
}

// Second build: listed wavefront k of the first build is built by wavefront k of this grid (as
// many as the first build's, so every listed one has its own), into a region of the spill pool
// exactly as wide as the first build found it needs.  The region is reserved with one atomic
// addition; when the pool is too small the wavefront stays unbuilt and the hold flag is raised,
// and q[1] (which sums what every listed wavefront reserved) is the size the host grows the pool
// to.  No wavefront loops here: every branch below is uniform across the wavefront (lane 0 alone
// does the atomics, then every lane reads its result), so no lane can be left behind while the
// others read lane 0.
__global__ __launch_bounds__(kBlock) void kNeighboursSpill(DeviceCtx c) {
    __shared__ WaveLds<1024> lds[kWavesPerBlock];
    if (stepHeld(c)) return;
    const int wave = threadIdx.x >> 6, lane = laneId();
    const uint32_t k = uniformU32(blockIdx.x * (uint32_t)kWavesPerBlock + (uint32_t)wave);
    const uint32_t listed =
        uniformU32((uint32_t)__hip_atomic_load(&c.nb.q[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    if (k >= listed) return;
    const uint32_t gw = uniformU32(c.nb.spill[k]);
    const uint32_t K = uniformU32(c.nb.waveNeed[gw]);
    const uint32_t UC = uniformU32((min(c.nb.unionLen[gw], 1u << 16) + 63u) & ~63u);
    const uint64_t listBytes = (128ull * K + 255ull) & ~255ull, bytes = listBytes + 4ull * UC;
    unsigned long long at = 0;
    if (lane == 0) at = atomicAdd(&c.nb.q[1], (unsigned long long)bytes);
    at = ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(at >> 32), 0) << 32) |
         (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)at, 0);
    if (at + bytes > c.nb.poolBytes) {
        if (lane == 0) {
            atomicOr(&c.cnt.v[3], 1ull);
            atomicOr(&c.cnt.v[kCntHold], 1ull);
        }
        return;
    }
    uint16_t* list = reinterpret_cast<uint16_t*>(c.nb.pool + at);
    uint32_t* U = reinterpret_cast<uint32_t*>(c.nb.pool + at + listBytes);
    const WaveBuild w = buildWave(c, lds[wave], gw, lane, list, K, U, UC);
    const bool fits = finishWave(c, gw, lane, w, K, UC);
    if (lane == 0) {
        c.nb.tab[gw] = WaveTable{list, U, K, UC};
        if (!fits) {  // K and UC are what the first build counted: a second walk finds the same set
            atomicOr(&c.cnt.v[3], 1ull);
            atomicOr(&c.cnt.v[kCntHold], 1ull);
        }
    }
}

__global__ void kNeighboursReset(DeviceCtx c) {
    if (threadIdx.x < 2) c.nb.q[threadIdx.x] = 0ull;
}

// The launch variants (waves per block, stack depth): (2, 1024), (4, 1024), (4, 512), (8, 512).
template <int W, int SCAP>
void buildAs(const DeviceCtx& c, hipStream_t s) {
    const uint32_t grid = (c.slots + (uint32_t)(W * kWave) - 1u) / (uint32_t)(W * kWave);
    kNeighbours<W, SCAP><<<grid, W * kWave, 0, s>>>(c);
}

}  // namespace

void launchNeighbours(const DeviceCtx& c, hipStream_t s) {
    kNeighboursReset<<<1, 64, 0, s>>>(c);
    const uint32_t w = c.tune.buildWaves, st = c.tune.buildStack;
    if (w == 2u) buildAs<2, 1024>(c, s);
    else if (w == 8u) buildAs<8, 512>(c, s);
    else if (st == 512u) buildAs<4, 512>(c, s);
    else buildAs<4, 1024>(c, s);
    kNeighboursSpill<<<c.blocks, kBlock, 0, s>>>(c);
}

}  // namespace astraea_hip
