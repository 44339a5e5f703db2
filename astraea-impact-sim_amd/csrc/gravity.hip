// gravity.hip — Barnes-Hut monopole gravity with a wavefront-cooperative traversal.
//
// Restates gravityBarnesHutKernel (engine/sph_gravity.cu:810-961):
//   leaf j != i:   a += -G m_j d / (r^2 + eps_ij^2 + 1e-20)^(3/2),  eps_ij = (eps_i + eps_j)/2,
//                  phi -= G m_j / sqrt(...), eps = soft_factor * max(h0, 1e-6)   (lines 842, 866-881)
//   node:          accepted when extent^2 < theta^2 r^2 and r^2 > 1e-20 (AABB-inside test when
//                  theta^2 > 1/3, line 855) for EVERY lane of the group; otherwise it is pushed
//                  for the whole group (lines 932-950); an accepted node uses eps_i only;
//                  a full stack forces acceptance (line 933).
// The group is the 64-lane wavefront (the reference's is the 32-lane warp), so the traversal
// state is uniform: the stack lives in LDS, the node index is read once with readfirstlane and
// the node and leaf records are uniform loads.  Only the per-lane opening test and the
// interaction arithmetic use vector ALUs.
#include "launch.hpp"

namespace astraea_hip {
namespace {

template <bool CHECK_INSIDE>
__global__ __launch_bounds__(kBlock) void kGravity(DeviceCtx c) {
    __shared__ int stackAll[kWavesPerBlock][kStack];
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
    const float theta2 = c.prm.theta2;
    const float soft = c.prm.softFactor;
    float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float epsI = 1.0e-6f;
    if (active) {
        pi = c.cur.pos[i];
        epsI = soft * fmaxf(c.cur.hu[i].y, 1.0e-6f);
    }
    const float epsI2 = epsI * epsI;
    float ax = 0.0f, ay = 0.0f, az = 0.0f, phi = 0.0f;
    uint32_t leafVisits = 0, nodeAccepts = 0, selfVisits = 0, forced = 0;

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
                const float epsJ = soft * fmaxf(c.cur.hu[j].y, 1.0e-6f);
                ++leafVisits;
                const uint32_t rel = j - waveBase;
                if (rel < 64u && ((am >> rel) & 1ull)) ++selfVisits;
                if (active && j != i) {
                    const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                    const float e = 0.5f * (epsI + epsJ);
                    const float d2 = dx * dx + dy * dy + dz * dz + e * e + 1.0e-20f;
                    const float ir = rsqrtf(d2);
                    const float gm = kGf * pj.w;
                    const float f = -gm * (ir * ir * ir);
                    ax += f * dx;
                    ay += f * dy;
                    az += f * dz;
                    phi -= gm * ir;
                }
            } else {
                const float4 cm = c.t.cm[cc];
                const float4 mn = c.t.bmin[cc];
                float dx, dy, dz, r2;
                bool accept;
                {
#pragma clang fp contract(off)
                    dx = pi.x - cm.x;
                    dy = pi.y - cm.y;
                    dz = pi.z - cm.z;
                    r2 = dist2f(dx, dy, dz);
                    const float lhs = mn.w * mn.w;
                    const float rhs = theta2 * r2;
                    accept = (lhs < rhs) && (r2 > 1.0e-20f);
                }
                if (CHECK_INSIDE && accept) {
                    const float4 mx = c.t.bmax[cc];
                    if (pi.x >= mn.x && pi.x <= mx.x && pi.y >= mn.y && pi.y <= mx.y && pi.z >= mn.z && pi.z <= mx.z)
                        accept = false;
                }
                const bool anyOpen = ballot(active && !accept) != 0ull;
                if (!anyOpen || sp >= kStack) {
                    ++nodeAccepts;
                    if (anyOpen) ++forced;
                    if (active) {
                        const float d2 = r2 + epsI2 + 1.0e-20f;
                        const float ir = rsqrtf(d2);
                        const float gm = kGf * cm.w;
                        const float f = -gm * (ir * ir * ir);
                        ax += f * dx;
                        ay += f * dy;
                        az += f * dz;
                        phi -= gm * ir;
                    }
                } else {
                    st[sp] = cc;
                    ++sp;
                }
            }
        }
    }
    if (active) c.d.grav[i] = make_float4(ax, ay, az, phi);
    if (lane == 0) {
        const unsigned long long lanes = (unsigned long long)__popcll(am);
        atomicAdd(&c.cnt.v[0], lanes * (unsigned long long)(leafVisits + nodeAccepts) - selfVisits);
        if (forced) atomicAdd(&c.cnt.v[1], (unsigned long long)forced);
    }
}

}  // namespace

void launchGravity(const DeviceCtx& c, hipStream_t s) {
    if (c.prm.theta2 > 0.333333f) kGravity<true><<<c.blocks, kBlock, 0, s>>>(c);
    else kGravity<false><<<c.blocks, kBlock, 0, s>>>(c);
}

}  // namespace astraea_hip
