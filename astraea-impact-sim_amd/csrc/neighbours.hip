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
// The tree walk is wave-cooperative: a node is opened when any lane's sphere meets its box.
#include "launch.hpp"

namespace astraea_hip {
namespace {

template <bool FILL>
__global__ __launch_bounds__(kBlock) void kNeighbours(DeviceCtx c) {
    __shared__ int stackAll[kWavesPerBlock][kStack];
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
        float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float Hb = 0.0f;
        if (active) {
            pi = c.cur.pos[i];
            const float4 hu = c.cur.hu[i];
            Hb = supportBound(hu.x, hu.y);
        }
        int* st = stackAll[wave];
        st[0] = (int)off;
        int sp = 1;
        while (sp > 0) {
            --sp;
            const int node = uniformInt(st[sp]);
            const int2 ch = c.t.child[node];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int cc = uniformInt(k == 0 ? ch.x : ch.y);
                if (cc < 0) {
                    const uint32_t j = (uint32_t)~cc;
                    const float4 pj = c.cur.pos[j];
                    const float4 hj = c.cur.hu[j];
                    const float R = fmaxf(Hb, supportBound(hj.x, hj.y));
                    const float r2 = dist2f(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z);
                    if (active && j != i && r2 < R * R) {
                        if (FILL && writable) c.nb.list[base + 64ull * cnt] = j;
                        ++cnt;
                    }
                } else {
                    const float4 mn = c.t.bmin[cc], mx = c.t.bmax[cc];
                    const float R = fmaxf(Hb, mx.w);
                    const bool ov = active && boxDist2f(pi.x, pi.y, pi.z, mn, mx) <= R * R;
                    if (ballot(ov) != 0ull) {
                        if (sp < kStack) {
                            st[sp] = cc;
                            ++sp;
                        } else {
                            stackFull = true;
                        }
                    }
                }
            }
        }
    }
    if (!FILL) {
        c.nb.count[i] = cnt;
        const uint32_t m = waveMaxU(cnt);
        if (lane == 0) c.nb.waveMax[gw] = m;
        if (lane == 0 && stackFull) atomicOr((unsigned long long*)&c.cnt.v[6], 1ull);
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
