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
// Layout: wavefront w owns a block at waveOff[w]; entry k of lane l is at
// waveOff[w] + 64 k + l, so a pass reading entry k on every lane issues one coalesced load.
// Build = count walk, exclusive scan of 64 * max count per wavefront, fill walk.
#include "launch.hpp"

namespace astraea_hip {
namespace {

constexpr int kSCap = 1024;  // per-wave node stack
constexpr int kLCap = 512;   // per-wave candidate-leaf list
constexpr int kFlushAt = kLCap - 256;

struct WaveLds {
    int stack[kSCap];
    int list[kLCap];
    float4 gp[kWave];  // group particle (x, y, z, support bound), w = -1 when inactive
};

// Wide walk: the wave expands up to 64 stack nodes per round (one per lane).  A node is opened
// when any group particle's sphere of radius max(Hb_p, node's max Hb) meets its box; the group's
// bounding sphere rejects most non-overlapping nodes without the per-particle test (a margin far
// above FP32 rounding keeps those decisions exact).  Leaves become candidates that every lane then
// tests against its own particle.
template <bool FILL>
__global__ __launch_bounds__(kBlock) void kNeighbours(DeviceCtx c) {
    __shared__ WaveLds lds[kWavesPerBlock];
    const uint32_t b = xcdBlock(blockIdx.x, gridDim.x);
    const int wave = threadIdx.x >> 6, lane = laneId();
    const uint32_t i = b * kBlock + threadIdx.x;
    const uint32_t gw = i >> 6;
    const uint32_t sg = uniformU32(c.seg.slotSeg[b * kBlock]);
    const uint32_t off = uniformU32(c.seg.off[sg]), n = uniformU32(c.seg.n[sg]);
    const bool active = i - off < n;
    uint32_t cnt = 0;
    bool stackFull = false;
    uint64_t base = 0;
    bool writable = true;
    if (FILL) {
        base = c.nb.waveOff[gw] + (uint64_t)lane;
        writable = c.nb.waveOff[gw] + 64ull * c.nb.waveMax[gw] <= c.nb.capacity;
    }
    if (ballot(active) != 0ull && n >= 2u) {
        WaveLds& L = lds[wave];
        float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float Hb = 0.0f;
        if (active) {
            pi = c.cur.pos[i];
            const float4 hu = c.cur.hu[i];
            Hb = supportBound(hu.x, hu.y);
        }
        L.gp[lane] = make_float4(pi.x, pi.y, pi.z, active ? Hb : -1.0f);
        float cx, cy, cz, R, HbMax;
        {
            const float lx = waveMinF(active ? pi.x : 3.0e38f), hx = waveMaxF(active ? pi.x : -3.0e38f);
            const float ly = waveMinF(active ? pi.y : 3.0e38f), hy = waveMaxF(active ? pi.y : -3.0e38f);
            const float lz = waveMinF(active ? pi.z : 3.0e38f), hz = waveMaxF(active ? pi.z : -3.0e38f);
            cx = 0.5f * (lx + hx); cy = 0.5f * (ly + hy); cz = 0.5f * (lz + hz);
            const float dx = pi.x - cx, dy = pi.y - cy, dz = pi.z - cz;
            R = waveMaxF(active ? sqrtf(dx * dx + dy * dy + dz * dz) : 0.0f);
            R = R * (1.0f + 1.0e-5f) + 1.0e-30f;
            HbMax = waveMaxF(active ? Hb : 0.0f);
        }
        const uint64_t lt = (1ull << lane) - 1ull;
        int sp = 1, nl = 0;
        L.stack[0] = (int)off;
        __builtin_amdgcn_wave_barrier();
        for (;;) {
            while (sp > 0 && nl <= kFlushAt) {
                const int k = sp < kWave ? sp : kWave;
                sp -= k;
                const bool has = lane < k;
                int2 ch = make_int2(0, 0);
                if (has) ch = c.t.child[L.stack[sp + lane]];
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int side = 0; side < 2; ++side) {
                    const int cc = side == 0 ? ch.x : ch.y;
                    const bool isLeaf = has && cc < 0;
                    bool open = false;
                    if (has && cc >= 0) {
                        const float4 mn = c.t.bmin[cc], mx = c.t.bmax[cc];
                        const float Rn = mx.w;
                        const float ddx = fmaxf(0.0f, fmaxf(mn.x - cx, cx - mx.x));
                        const float ddy = fmaxf(0.0f, fmaxf(mn.y - cy, cy - mx.y));
                        const float ddz = fmaxf(0.0f, fmaxf(mn.z - cz, cz - mx.z));
                        const float db = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
                        const float reach = fmaxf(HbMax, Rn);
                        const float slack = 2.0e-6f * (db + R + reach) + 1.0e-30f;
                        if (db - R - slack <= reach) {
                            for (int j = 0; j < kWave && !open; ++j) {
                                const float4 g = L.gp[j];
                                if (g.w < 0.0f) continue;
                                const float Rr = fmaxf(g.w, Rn);
                                open = boxDist2f(g.x, g.y, g.z, mn, mx) <= Rr * Rr;
                            }
                        }
                    }
                    const uint64_t mPush = ballot(open);
                    int pushAt = sp + __popcll(mPush & lt);
                    if (open && pushAt >= kSCap) {
                        open = false;
                        stackFull = true;
                    }
                    if (open) L.stack[pushAt] = cc;
                    const int nPush = __popcll(mPush);
                    sp = sp + nPush < kSCap ? sp + nPush : kSCap;
                    const uint64_t mList = ballot(isLeaf);
                    if (isLeaf) L.list[nl + __popcll(mList & lt)] = ~cc;
                    nl += __popcll(mList);
                }
                __builtin_amdgcn_wave_barrier();
            }
            // candidates: every lane tests its own particle, four records in flight
            int e = 0;
            for (; e + 4 <= nl; e += 4) {
                uint32_t j[4];
                float4 pj[4], hj[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    j[u] = (uint32_t)uniformInt(L.list[e + u]);
                    pj[u] = c.cur.pos[j[u]];
                    hj[u] = c.cur.hu[j[u]];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const float Rr = fmaxf(Hb, supportBound(hj[u].x, hj[u].y));
                    const float r2 = dist2f(pi.x - pj[u].x, pi.y - pj[u].y, pi.z - pj[u].z);
                    if (active && j[u] != i && r2 < Rr * Rr) {
                        if (FILL && writable) c.nb.list[base + 64ull * cnt] = j[u];
                        ++cnt;
                    }
                }
            }
            for (; e < nl; ++e) {
                const uint32_t jj = (uint32_t)uniformInt(L.list[e]);
                const float4 pj = c.cur.pos[jj];
                const float4 hj = c.cur.hu[jj];
                const float Rr = fmaxf(Hb, supportBound(hj.x, hj.y));
                const float r2 = dist2f(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z);
                if (active && jj != i && r2 < Rr * Rr) {
                    if (FILL && writable) c.nb.list[base + 64ull * cnt] = jj;
                    ++cnt;
                }
            }
            nl = 0;
            __builtin_amdgcn_wave_barrier();
            if (sp == 0) break;
        }
    }
    if (!FILL) {
        c.nb.count[i] = cnt;
        const uint32_t m = waveMaxU(cnt);
        if (lane == 0) c.nb.waveMax[gw] = m;
        if (ballot(stackFull) != 0ull && lane == 0) atomicOr((unsigned long long*)&c.cnt.v[6], 1ull);
    }
}

// Exclusive scan of 64 * waveMax over all wavefronts (one block of 1024 threads).
__global__ __launch_bounds__(1024) void kNeighbourScan(DeviceCtx c, uint32_t waves) {
    __shared__ unsigned long long part[16];
    const uint32_t t = threadIdx.x, lane = laneId(), w = t >> 6;
    const uint32_t per = (waves + 1023u) / 1024u;
    const uint32_t s0 = t * per, s1 = min(waves, s0 + per);
    unsigned long long sum = 0;
    for (uint32_t k = s0; k < s1; ++k) sum += 64ull * c.nb.waveMax[k];
    unsigned long long inc = sum;
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long v = __shfl_up(inc, o, 64);
        if ((int)lane >= o) inc += v;
    }
    if (lane == 63) part[w] = inc;
    __syncthreads();
    unsigned long long pre = 0, tot = 0;
    for (uint32_t q = 0; q < 16; ++q) {
        if (q < w) pre += part[q];
        tot += part[q];
    }
    unsigned long long run = pre + inc - sum;
    for (uint32_t k = s0; k < s1; ++k) {
        c.nb.waveOff[k] = run;
        run += 64ull * c.nb.waveMax[k];
    }
    if (t == 0) {
        c.cnt.v[2] = tot;
        if (tot > c.nb.capacity) c.cnt.v[3] = 1ull;
    }
}

}  // namespace

void launchNeighbours(const DeviceCtx& c, hipStream_t s) {
    kNeighbours<false><<<c.blocks, kBlock, 0, s>>>(c);
    kNeighbourScan<<<1, 1024, 0, s>>>(c, c.slots / kWave);
    kNeighbours<true><<<c.blocks, kBlock, 0, s>>>(c);
}

}  // namespace astraea_hip
