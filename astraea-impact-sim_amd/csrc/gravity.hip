// gravity.hip — Barnes-Hut monopole gravity, wavefront-cooperative, wide traversal.
//
// Restates gravityBarnesHutKernel (engine/sph_gravity.cu:810-961):
//   leaf j != i:   a += -G m_j d / (r^2 + eps_ij^2 + 1e-20)^(3/2),  eps_ij = (eps_i + eps_j)/2,
//                  phi -= G m_j / sqrt(...), eps = soft_factor * max(h0, 1e-6)   (lines 842, 866-881)
//   node:          accepted when extent^2 < theta^2 r^2 and r^2 > 1e-20 (AABB-inside test when
//                  theta^2 > 1/3, line 855) for EVERY particle of the group; otherwise it is opened
//                  for the whole group (lines 932-950); an accepted node uses eps_i only.
// The group is the 64-lane wavefront (the reference's is the 32-lane warp).
//
// The reference pops one node per iteration, so each wave waits on two dependent loads per node.
// Here the wave expands up to 64 nodes of its LDS stack at once, one per lane: each lane loads its
// node's children and decides the group opening test for them.  Most decisions come from the
// group's bounding sphere (centre c, radius R): every particle's distance to a node's centre of
// mass lies in [d - R, d + R], d = |c - com|, and with a relative margin far above FP32 rounding
// the per-particle FP32 test is then known to accept (or reject) for all particles; only the
// remaining nodes run the exact per-particle test, reading the group's positions from LDS.  The
// interaction set is exactly the reference's.  Accepted nodes and leaves go to an LDS list that
// the whole wave then evaluates, four records in flight at a time.
#include "launch.hpp"

namespace astraea_hip {
namespace {

constexpr int kSCap = 1024;  // per-wave node stack
constexpr int kLCap = 512;   // per-wave interaction list
constexpr int kFlushAt = kLCap - 256;

struct WaveLds {
    int stack[kSCap];
    int list[kLCap];
    float4 gp[kWave];  // group particle positions, w = 1 when active
};

__device__ __forceinline__ void interact(float px, float py, float pz, float4 q, float e2, float& ax, float& ay, float& az,
                                         float& phi) {
    const float dx = px - q.x, dy = py - q.y, dz = pz - q.z;
    const float d2 = dx * dx + dy * dy + dz * dz + e2 + 1.0e-20f;
    const float ir = rsqrtf(d2);
    const float gm = kGf * q.w;
    const float f = -gm * (ir * ir * ir);
    ax += f * dx;
    ay += f * dy;
    az += f * dz;
    phi -= gm * ir;
}

template <bool CHECK_INSIDE>
__global__ __launch_bounds__(kBlock) void kGravity(DeviceCtx c) {
    __shared__ WaveLds lds[kWavesPerBlock];
    const uint32_t b = xcdBlock(blockIdx.x, gridDim.x);
    const int wave = threadIdx.x >> 6, lane = laneId();
    const uint32_t i = b * kBlock + threadIdx.x;
    const uint32_t waveBase = b * kBlock + (uint32_t)wave * kWave;
    const uint32_t sg = uniformU32(c.seg.slotSeg[b * kBlock]);
    const uint32_t off = uniformU32(c.seg.off[sg]), n = uniformU32(c.seg.n[sg]);
    const bool active = i - off < n;
    const uint64_t am = ballot(active);
    if (am == 0ull) return;
    if (n == 1u) {
        if (active) c.d.grav[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    WaveLds& L = lds[wave];
    const float theta2 = c.prm.theta2;
    const float theta = sqrtf(theta2);
    const float soft = c.prm.softFactor;
    float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float epsI = 1.0e-6f;
    if (active) {
        pi = c.cur.pos[i];
        epsI = soft * fmaxf(c.cur.hu[i].y, 1.0e-6f);
    }
    const float epsI2 = epsI * epsI;
    L.gp[lane] = make_float4(pi.x, pi.y, pi.z, active ? 1.0f : 0.0f);

    // group bounding sphere over the active particles (centre of the box, radius of the farthest)
    float cx, cy, cz, R;
    {
        const float lx = waveMinF(active ? pi.x : 3.0e38f), hx = waveMaxF(active ? pi.x : -3.0e38f);
        const float ly = waveMinF(active ? pi.y : 3.0e38f), hy = waveMaxF(active ? pi.y : -3.0e38f);
        const float lz = waveMinF(active ? pi.z : 3.0e38f), hz = waveMaxF(active ? pi.z : -3.0e38f);
        cx = 0.5f * (lx + hx); cy = 0.5f * (ly + hy); cz = 0.5f * (lz + hz);
        const float dx = pi.x - cx, dy = pi.y - cy, dz = pi.z - cz;
        R = waveMaxF(active ? sqrtf(dx * dx + dy * dy + dz * dz) : 0.0f);
        R = R * (1.0f + 1.0e-5f) + 1.0e-30f;
    }

    float ax = 0.0f, ay = 0.0f, az = 0.0f, phi = 0.0f;
    uint32_t listed = 0, selfVisits = 0, forced = 0;
    const uint64_t lt = (1ull << lane) - 1ull;

    int sp = 1, nl = 0;
    L.stack[0] = (int)off;
    __builtin_amdgcn_wave_barrier();
    for (;;) {
        // ---- expand: up to 64 stack nodes per round, one per lane ----
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
                const bool isInt = has && cc >= 0;
                bool accept = false;
                if (isInt) {
                    const float4 cm = c.t.cm[cc];
                    const float4 mn = c.t.bmin[cc];
                    const float ext = mn.w;
                    const float ddx = cx - cm.x, ddy = cy - cm.y, ddz = cz - cm.z;
                    const float d = sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
                    const float slack = 2.0e-6f * (d + R) + 1.0e-30f;
                    const float dLo = d - R - slack, dHi = d + R + slack;
                    int decided = 0;  // 1 accept, 2 reject, 0 exact test
                    if (!CHECK_INSIDE && dLo > 1.0e-9f && theta * dLo > ext * (1.0f + 1.0e-5f)) decided = 1;
                    else if (theta * dHi < ext * (1.0f - 1.0e-5f)) decided = 2;
                    if (decided == 1) {
                        accept = true;
                    } else if (decided == 0) {
                        float4 mx = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        if (CHECK_INSIDE) mx = c.t.bmax[cc];
                        accept = true;
                        for (int j = 0; j < kWave && accept; ++j) {
                            const float4 g = L.gp[j];
                            if (g.w == 0.0f) continue;
                            bool a;
                            {
#pragma clang fp contract(off)
                                const float r2 = dist2f(g.x - cm.x, g.y - cm.y, g.z - cm.z);
                                const float lhs = ext * ext;
                                const float rhs = theta2 * r2;
                                a = (lhs < rhs) && (r2 > 1.0e-20f);
                            }
                            if (CHECK_INSIDE && a && g.x >= mn.x && g.x <= mx.x && g.y >= mn.y && g.y <= mx.y &&
                                g.z >= mn.z && g.z <= mx.z)
                                a = false;
                            accept = a;
                        }
                    }
                }
                const bool toList = isLeaf || (isInt && accept);
                bool toPush = isInt && !accept;
                const uint64_t mPush = ballot(toPush);
                const int nPush = __popcll(mPush);
                int pushAt = sp + __popcll(mPush & lt);
                if (toPush && pushAt >= kSCap) toPush = false;  // forced acceptance (reference: full stack)
                const bool forcedHere = isInt && !accept && !toPush;
                const bool listIt = toList || forcedHere;
                const uint64_t mList = ballot(listIt);
                if (listIt) L.list[nl + __popcll(mList & lt)] = cc;
                if (toPush) L.stack[pushAt] = cc;
                nl += __popcll(mList);
                sp = sp + nPush < kSCap ? sp + nPush : kSCap;
                forced += (uint32_t)__popcll(ballot(forcedHere));
                if (isLeaf) {
                    const uint32_t rel = (uint32_t)(~cc) - waveBase;
                    if (rel < 64u && ((am >> rel) & 1ull)) selfVisits += 1;  // counted per lane, summed below
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
        // ---- evaluate the list: four records in flight ----
        listed += (uint32_t)nl;
        int e = 0;
        for (; e + 4 <= nl; e += 4) {
            int id[4];
            float4 q[4];
            float ej[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                id[u] = uniformInt(L.list[e + u]);
                if (id[u] < 0) {
                    q[u] = c.cur.pos[~id[u]];
                    ej[u] = c.cur.hu[~id[u]].y;
                } else {
                    q[u] = c.t.cm[id[u]];
                    ej[u] = 0.0f;
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (id[u] < 0) {
                    const float eij = 0.5f * (epsI + soft * fmaxf(ej[u], 1.0e-6f));
                    if (active && (uint32_t)(~id[u]) != i) interact(pi.x, pi.y, pi.z, q[u], eij * eij, ax, ay, az, phi);
                } else if (active) {
                    interact(pi.x, pi.y, pi.z, q[u], epsI2, ax, ay, az, phi);
                }
            }
        }
        for (; e < nl; ++e) {
            const int id = uniformInt(L.list[e]);
            if (id < 0) {
                const uint32_t j = (uint32_t)~id;
                const float4 q = c.cur.pos[j];
                const float eij = 0.5f * (epsI + soft * fmaxf(c.cur.hu[j].y, 1.0e-6f));
                if (active && j != i) interact(pi.x, pi.y, pi.z, q, eij * eij, ax, ay, az, phi);
            } else if (active) {
                interact(pi.x, pi.y, pi.z, c.t.cm[id], epsI2, ax, ay, az, phi);
            }
        }
        nl = 0;
        __builtin_amdgcn_wave_barrier();
        if (sp == 0) break;
    }
    if (active) c.d.grav[i] = make_float4(ax, ay, az, phi);
    // interactions = active lanes x listed entries - each active particle's own leaf
    uint32_t selfSum = selfVisits;
    for (int o = 32; o > 0; o >>= 1) selfSum += (uint32_t)__shfl_xor((int)selfSum, o, 64);
    if (lane == 0) {
        const unsigned long long lanes = (unsigned long long)__popcll(am);
        atomicAdd(&c.cnt.v[0], lanes * (unsigned long long)listed - selfSum);
        if (forced) atomicAdd(&c.cnt.v[1], (unsigned long long)forced);
    }
}

}  // namespace

void launchGravity(const DeviceCtx& c, hipStream_t s) {
    if (c.prm.theta2 > 0.333333f) kGravity<true><<<c.blocks, kBlock, 0, s>>>(c);
    else kGravity<false><<<c.blocks, kBlock, 0, s>>>(c);
}

}  // namespace astraea_hip
