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
// Here the wave expands up to 32 stack nodes of its LDS stack at once, a lane per child: each
// lane loads its child and decides the group opening test for it.  The per-particle FP32 test is rewritten as r^2 >= T with T the exact
// FP32 acceptance threshold of the node, and most decisions come from the squared distances
// between the node's centre of mass and the group's bounding box (nearest point and farthest
// corner): with a relative margin far above FP32 rounding the test is then known to pass (or
// fail) for all particles.  Of the remaining nodes,
// those that one representative particle fails (the nearest sub-group's extreme particle toward
// the node, tested with the same FP32 expression) are opened at once; the rest are tested
// exactly, one node at a time across the lanes (lane = particle), read back from LDS.
// With the reference's groups of 32 the interaction set is exactly the reference's as long as no
// node is force-accepted (forced acceptances counted zero): this stack holds 1024 entries and
// fills breadth-first, the reference's 192 depth-first (sph_gravity.cu:22, 933), so when a stack
// fills the two accept different nodes.
//
// Accepted nodes and leaves are staged in LDS as ready records (x, y, z, G m), loaded by their
// lanes in parallel; the wave then evaluates them with broadcast LDS reads.  A particle's own
// leaf is evaluated like the others (zero separation: no force) and its potential term is taken
// back at the end, which keeps the per-pair test out of the inner loop.
#include "launch.hpp"

namespace astraea_hip {
namespace {

constexpr int kSub = 4;                     // sub-groups of a wavefront with boxes of their own
constexpr int kSubLanes = kWave / kSub;
// The staged records are evaluated once either list holds more than kFlushAt of them, checked
// after every round, so a list never exceeds kFlushAt + 64 (a round lists at most a record per
// lane).  The undecided nodes of a round are staged past the kFlushAt records a list can hold
// when the round starts.  The LDS a wavefront holds sets how many share a CU.
constexpr int kFlushAt = 16;
constexpr int kRCap = kFlushAt + kWave;
constexpr int kPend = kFlushAt;

template <int SCAP>  // per-wave node stack: 1024 or 512 entries (a launch variant)
struct WaveLds {
    int stack[SCAP];
    float4 node[kRCap];     // accepted node: (com x, y, z, G m)
    float4 leaf[kRCap];     // leaf: (x, y, z, G m)
    float leafEps[kRCap];   // eps_j / 2
    uint32_t self[kWave];   // lane l's own leaf was listed
};

// Least float T with fl(theta2 * T) > e2 and T > 1e-20: for every float r2,
// (e2 < fl(theta2 * r2)) && (r2 > 1e-20)  <=>  r2 >= T, because rounding is monotone.
// Starts within a few ulp of e2 / theta2 and steps to the exact boundary.
__device__ __forceinline__ float acceptThreshold(float e2, float theta2, float invTheta2, bool pow2) {
    if (!(theta2 > 0.0f)) return __builtin_inff();
    const float floor20 = __uint_as_float(__float_as_uint(1.0e-20f) + 1u);
    float x = e2 * invTheta2;
    if (pow2 && x < 3.0e38f) {
        // theta^2 * x is exact, so the boundary is e2 / theta^2 itself and T is the next float
        return fmaxf(__uint_as_float(__float_as_uint(x) + 1u), floor20);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float d = __uint_as_float(__float_as_uint(x) - 1u);
        if (x > 0.0f && theta2 * d > e2) x = d;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (!(theta2 * x > e2)) x = __uint_as_float(__float_as_uint(x) + 1u);
    return fmaxf(x, floor20);
}

// GROUP particles share the opening decisions of one wavefront: 64 (all lanes), or 32, the
// reference's warp groups, decided by lanes 0-31 while lanes 32-63 help with their partners'
// exact tests and interactions.  A block is W wavefronts over W * GROUP slots, and a wavefront's
// GROUP slots never straddle a segment (segments start at multiples of 256).  W and the stack
// depth SCAP are launch variants: they set how many wavefronts share a block's LDS and a CU's L1
// (neighbouring groups walk the same nodes), never the results while the stack does not fill.
template <bool CHECK_INSIDE, int GROUP, int W, int SCAP>
__global__ __launch_bounds__(W * kWave) void kGravity(DeviceCtx c) {
    static_assert(GROUP == kWave || GROUP == kWave / 2, "a group is a wavefront or half of one");
    constexpr int kSubG = kSub * GROUP / kWave;  // the sub-groups that hold the group's lanes
    constexpr uint32_t kSlots = (uint32_t)(W * GROUP);
    __shared__ WaveLds<SCAP> lds[W];
    if (stepHeld(c)) return;
    const uint32_t b = xcdBlock(blockIdx.x, gridDim.x);
    const int wave = threadIdx.x >> 6, lane = laneId();
    const uint32_t waveBase = b * kSlots + (uint32_t)wave * GROUP;
    if (waveBase >= c.slots) return;  // the last block of a grid over a multiple of 256 slots
    const uint32_t i = waveBase + (uint32_t)lane;
    const uint32_t sg = uniformU32(c.seg.slotSeg[waveBase]);
    const uint32_t off = uniformU32(c.seg.off[sg]), n = uniformU32(c.seg.n[sg]);
    const bool active = lane < GROUP && i - off < n;
    const uint64_t am = ballot(active);
    if (am == 0ull) return;
    if (n == 1u) {
        if (active) c.d.grav[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    WaveLds<SCAP>& L = lds[wave];
    const float theta2 = c.prm.theta2;
    const float invTheta2 = theta2 > 0.0f ? 1.0f / theta2 : 0.0f;
    const bool pow2 = c.prm.theta2Pow2 != 0u;
    const float soft = c.prm.softFactor;
    // the particle a lane evaluates: its own; at GROUP 32, lanes 32-63 take their partner's
    // (lane - 32) and every other record, and the two partial sums are added at the end.  The
    // walk's decisions use the group's lanes only (`active`, `am`).
    const uint32_t ie = waveBase + (uint32_t)(lane % GROUP);
    float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float epsI = 1.0e-6f;
    if (ie - off < n) {
        pi = c.cur.pos[ie];
        epsI = soft * fmaxf(c.cur.hu[ie].y, 1.0e-6f);
    }
    const float halfEpsI = 0.5f * epsI;
    const float e2Node = __builtin_fmaf(epsI, epsI, 1.0e-20f);
    L.self[lane] = 0u;

    // bounding boxes of the kSubG sub-groups of kSubLanes lanes over their active particles (a
    // sub-group without one gets an inverted box: every node is apart from it and accepted by it)
    float bl[kSub][3], bh[kSub][3];
    {
        float v[6] = {active ? pi.x : 3.0e38f, active ? pi.y : 3.0e38f, active ? pi.z : 3.0e38f,
                      active ? pi.x : -3.0e38f, active ? pi.y : -3.0e38f, active ? pi.z : -3.0e38f};
        for (int o = 1; o < kSubLanes; o <<= 1) {
            for (int a = 0; a < 3; ++a) v[a] = fminf(v[a], __shfl_xor(v[a], o, 64));
            for (int a = 3; a < 6; ++a) v[a] = fmaxf(v[a], __shfl_xor(v[a], o, 64));
        }
#pragma unroll
        for (int g = 0; g < kSubG; ++g) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                bl[g][a] = readLaneF(v[a], kSubLanes * g);
                bh[g][a] = readLaneF(v[3 + a], kSubLanes * g);
            }
        }
    }

    // per sub-group, the lanes of its extreme active particles along -x, +x, -y, +y, -z, +z (one
    // byte each, 0xff for an empty sub-group) and its box centre
    uint64_t ext[kSub];
    float cen[kSub][3];
#pragma unroll
    for (int g = 0; g < kSubG; ++g) {
        const uint64_t gm = am & (0xffffull << (kSubLanes * g));
        uint64_t packed = 0ull;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float x = a == 0 ? pi.x : (a == 1 ? pi.y : pi.z);
            const uint64_t lo = ballot(x == bl[g][a]) & gm, hi = ballot(x == bh[g][a]) & gm;
            packed |= (uint64_t)(lo ? (uint32_t)__builtin_ctzll(lo) : 0xffu) << (16 * a);
            packed |= (uint64_t)(hi ? (uint32_t)__builtin_ctzll(hi) : 0xffu) << (16 * a + 8);
            cen[g][a] = 0.5f * bl[g][a] + 0.5f * bh[g][a];
        }
        ext[g] = packed;
    }

    float ax = 0.0f, ay = 0.0f, az = 0.0f, phi = 0.0f;
    uint32_t listed = 0, forced = 0;
    const uint64_t lt = (1ull << lane) - 1ull;
    constexpr int kStride = kWave / GROUP;  // records a lane steps over in the evaluation
    const int e0 = GROUP == kWave ? 0 : lane / GROUP;

    int sp = 1, nn = 0, nf = 0;
    // evaluates the staged records and empties both lists
    const auto evaluate = [&]() {
        __builtin_amdgcn_wave_barrier();
        listed += (uint32_t)(nn + nf);
#pragma unroll 4
        for (int e = e0; e < nn; e += kStride) {
            const float4 q = L.node[e];
            const float dx = pi.x - q.x, dy = pi.y - q.y, dz = pi.z - q.z;
            const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, e2Node)));
            const float ir = __builtin_amdgcn_rsqf(d2);  // d2 >= 1e-20: never denormal
            const float t = q.w * ir;
            const float f = t * (ir * ir);
            ax = __builtin_fmaf(-f, dx, ax);
            ay = __builtin_fmaf(-f, dy, ay);
            az = __builtin_fmaf(-f, dz, az);
            phi -= t;
        }
#pragma unroll 4
        for (int e = e0; e < nf; e += kStride) {
            const float4 q = L.leaf[e];
            const float eij = halfEpsI + L.leafEps[e];
            const float dx = pi.x - q.x, dy = pi.y - q.y, dz = pi.z - q.z;
            const float e2 = __builtin_fmaf(eij, eij, 1.0e-20f);
            const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, e2)));
            const float ir = __builtin_amdgcn_rsqf(d2);
            const float t = q.w * ir;
            const float f = t * (ir * ir);
            ax = __builtin_fmaf(-f, dx, ax);
            ay = __builtin_fmaf(-f, dy, ay);
            az = __builtin_fmaf(-f, dz, az);
            phi -= t;
        }
        nn = 0;
        nf = 0;
        __builtin_amdgcn_wave_barrier();
    };
    L.stack[0] = (int)off;
    __builtin_amdgcn_wave_barrier();
    for (;;) {
        // ---- expand: up to 32 stack nodes per round, a lane per child ----
        while (sp > 0 && nn <= kFlushAt && nf <= kFlushAt) {
            // lanes 2t and 2t + 1 take node t's two children.  At most 32 nodes a round keeps
            // every round to one pass over full lanes (rounds of 33-64 nodes, a lane per node and
            // its children in two passes, left lanes idle: 8.2 against 7.8 ms on the sweep).
            const int k = sp < kWave / 2 ? sp : kWave / 2;
            sp -= k;
            const int at = lane >> 1;
            const bool has = at < k;
            int2 ch = make_int2(0, 0);
            if (has) ch = c.t.child[L.stack[sp + at]];
            __builtin_amdgcn_wave_barrier();
            {
                const int cc = (lane & 1) == 0 ? ch.x : ch.y;
                const bool isLeaf = has && cc < 0;
                const bool isInt = has && cc >= 0;
                float4 cm = make_float4(0.0f, 0.0f, 0.0f, 0.0f), mn = cm, mx = cm;
                float4 rec = cm;
                float leafHalfEps = 0.0f, T = 0.0f;
                bool accept = false, undecided = false;
                int repG = 0;
                if (isInt) {
                    cm = c.t.cm[cc];
                    mn = c.t.bmin[cc];
                    if (CHECK_INSIDE) mx = c.t.bmax[cc];
                    {
#pragma clang fp contract(off)
                        const float e2 = mn.w * mn.w;
                        T = acceptThreshold(e2, theta2, invTheta2, pow2);
                    }
                    // per sub-group: squared distance from the com to its box (g2) and to the
                    // box's farthest corner (f2); every particle's FP32 r^2 lies within 1e-6
                    // (relative) of [g2, f2], far inside the 4e-5 margins below.  The node is
                    // accepted when every sub-group surely accepts it, opened when one sub-group
                    // surely rejects it (all its particles fail), and tested exactly otherwise.
                    bool allPass = true, oneFails = false;
                    float g2min = 3.0e38f;
                    int gNear = 0;
#pragma unroll
                    for (int g = 0; g < kSubG; ++g) {
                        const float ux = bl[g][0] - cm.x, vx = cm.x - bh[g][0];
                        const float uy = bl[g][1] - cm.y, vy = cm.y - bh[g][1];
                        const float uz = bl[g][2] - cm.z, vz = cm.z - bh[g][2];
                        const float gx = fmaxf(0.0f, fmaxf(ux, vx)), fx = fmaxf(fabsf(ux), fabsf(vx));
                        const float gy = fmaxf(0.0f, fmaxf(uy, vy)), fy = fmaxf(fabsf(uy), fabsf(vy));
                        const float gz = fmaxf(0.0f, fmaxf(uz, vz)), fz = fmaxf(fabsf(uz), fabsf(vz));
                        const float g2 = gx * gx + gy * gy + gz * gz;
                        const float f2 = fx * fx + fy * fy + fz * fz;
                        bool apart = true;  // no particle of the sub-group can lie inside the node's box
                        if (CHECK_INSIDE)
                            apart = bh[g][0] < mn.x || bl[g][0] > mx.x || bh[g][1] < mn.y || bl[g][1] > mx.y ||
                                    bh[g][2] < mn.z || bl[g][2] > mx.z;
                        allPass = allPass && apart && g2 > T * (1.0f + 4.0e-5f);
                        oneFails = oneFails || f2 * (1.0f + 4.0e-5f) < T;
                        if (g2 < g2min) {
                            g2min = g2;
                            gNear = g;
                        }
                    }
                    if (allPass) accept = true;
                    else if (!oneFails) undecided = true;
                    rec = make_float4(cm.x, cm.y, cm.z, kGf * cm.w);
                    repG = gNear;
                } else if (isLeaf) {
                    const uint32_t j = (uint32_t)~cc;
                    const float4 pj = c.cur.pos[j];
                    rec = make_float4(pj.x, pj.y, pj.z, kGf * pj.w);
                    leafHalfEps = 0.5f * (soft * fmaxf(c.cur.hu[j].y, 1.0e-6f));
                    const uint32_t rel = j - waveBase;
                    if (rel < 64u && ((am >> rel) & 1ull)) L.self[rel] = 1u;
                }
                // representative test: the group opens a node as soon as one of its particles fails
                // the opening test, so the extreme particle toward the node of the nearest sub-group
                // (the likeliest to fail) is tested first, with the exact test's own FP32
                // expression; a failure decides the node (opened) without the group test.  On the
                // 64-scenario sweep this settles 56 % of the undecided nodes.  The particle's
                // coordinates are read from its lane, with every lane active.
                {
                    uint64_t packed = ext[0];
                    float c0 = cen[0][0], c1 = cen[0][1], c2 = cen[0][2];
#pragma unroll
                    for (int h = 1; h < kSubG; ++h) {
                        const bool is = repG == h;
                        packed = is ? ext[h] : packed;
                        c0 = is ? cen[h][0] : c0;
                        c1 = is ? cen[h][1] : c1;
                        c2 = is ? cen[h][2] : c2;
                    }
                    const float dx = cm.x - c0, dy = cm.y - c1, dz = cm.z - c2;
                    const float adx = fabsf(dx), ady = fabsf(dy), adz = fabsf(dz);
                    const uint32_t sel = adx >= ady && adx >= adz ? (dx > 0.0f ? 1u : 0u)
                                         : (ady >= adz ? (dy > 0.0f ? 3u : 2u) : (dz > 0.0f ? 5u : 4u));
                    const uint32_t rl = (uint32_t)(packed >> (8u * sel)) & 0xffu;
                    const int src = rl < 64u ? (int)rl : lane;
                    const float rx = __shfl(pi.x, src, 64), ry = __shfl(pi.y, src, 64), rz = __shfl(pi.z, src, 64);
                    bool repFail;
                    {
#pragma clang fp contract(off)
                        const float r2 = dist2f(rx - cm.x, ry - cm.y, rz - cm.z);
                        repFail = rl < 64u && !(r2 >= T);
                    }
                    if (repFail) undecided = false;  // opened (accept stays false)
                }
                // exact group test of the undecided nodes, one node at a time across the lanes
                // (lane = particle); the nodes are staged in the free tail of the node list
                const uint64_t um = ballot(undecided);
                const int myPend = __popcll(um & lt);
                if (undecided) L.node[kPend + myPend] = make_float4(cm.x, cm.y, cm.z, T);
                __builtin_amdgcn_wave_barrier();
                uint64_t okMask = 0ull;
                if (CHECK_INSIDE) {
                    uint64_t rest = um;
                    for (int u = 0; rest != 0ull; ++u) {
                        const int src = (int)__builtin_ctzll(rest);
                        rest &= rest - 1ull;
                        const float4 q = L.node[kPend + u];
                        bool fail;
                        {
#pragma clang fp contract(off)
                            const float r2 = dist2f(pi.x - q.x, pi.y - q.y, pi.z - q.z);
                            fail = !(r2 >= q.w);
                        }
                        if (!fail) {
                            const float x0 = readLaneF(mn.x, src), y0 = readLaneF(mn.y, src), z0 = readLaneF(mn.z, src);
                            const float x1 = readLaneF(mx.x, src), y1 = readLaneF(mx.y, src), z1 = readLaneF(mx.z, src);
                            fail = pi.x >= x0 && pi.x <= x1 && pi.y >= y0 && pi.y <= y1 && pi.z >= z0 && pi.z <= z1;
                        }
                        if ((ballot(fail) & am) == 0ull) okMask |= 1ull << u;
                    }
                } else {
                    // four nodes per iteration (eight at GROUP 32: lanes 32-63 test the odd ones
                    // for their partners' particles): their LDS reads overlap; bits past the
                    // pending count are never read
                    const int nu = __popcll(um);
                    for (int u = 0; u < nu; u += 4 * kStride) {
                        float4 q[4];
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const int at = kPend + u + kStride * v + e0;  // slots past nu hold stale records
                            q[v] = L.node[at < kRCap ? at : kRCap - 1];
                        }
                        uint32_t bits = 0u;
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            bool fail;
                            {
#pragma clang fp contract(off)
                                const float r2 = dist2f(pi.x - q[v].x, pi.y - q[v].y, pi.z - q[v].z);
                                fail = !(r2 >= q[v].w);
                            }
                            const uint64_t b = ballot(fail);
                            if constexpr (GROUP == kWave) {
                                bits |= ((b & am) == 0ull ? 1u : 0u) << v;
                            } else {
                                bits |= ((b & am) == 0ull ? 1u : 0u) << (2 * v);
                                bits |= (((b >> GROUP) & am) == 0ull ? 1u : 0u) << (2 * v + 1);
                            }
                        }
                        okMask |= (uint64_t)bits << u;
                    }
                }
                if (undecided) accept = ((okMask >> myPend) & 1ull) != 0ull;
                __builtin_amdgcn_wave_barrier();
                bool toPush = isInt && !accept;
                const uint64_t mPush = ballot(toPush);
                const int nPush = __popcll(mPush);
                const int pushAt = sp + __popcll(mPush & lt);
                if (toPush && pushAt >= SCAP) toPush = false;  // forced acceptance (reference: full stack)
                const bool forcedHere = isInt && !accept && !toPush;
                const bool asNode = isInt && (accept || forcedHere);
                const uint64_t mNode = ballot(asNode), mLeaf = ballot(isLeaf);
                if (asNode) L.node[nn + __popcll(mNode & lt)] = rec;
                if (isLeaf) {
                    const int at = nf + __popcll(mLeaf & lt);
                    L.leaf[at] = rec;
                    L.leafEps[at] = leafHalfEps;
                }
                if (toPush) L.stack[pushAt] = cc;
                nn += __popcll(mNode);
                nf += __popcll(mLeaf);
                sp = sp + nPush < SCAP ? sp + nPush : SCAP;
                forced += (uint32_t)__popcll(ballot(forcedHere));
            }
            __builtin_amdgcn_wave_barrier();
        }
        evaluate();
        if (sp == 0) break;
    }
    // the particle's own leaf: zero separation, so only its potential term is taken back
    // (the same expression as in the loop: eps_ii = eps_i)
    if (GROUP < kWave) {
        ax += __shfl_xor(ax, GROUP, kWave);
        ay += __shfl_xor(ay, GROUP, kWave);
        az += __shfl_xor(az, GROUP, kWave);
        phi += __shfl_xor(phi, GROUP, kWave);
    }
    const bool selfListed = L.self[lane] != 0u;
    if (active && selfListed) phi += (kGf * pi.w) * __builtin_amdgcn_rsqf(__builtin_fmaf(epsI, epsI, 1.0e-20f));
    if (active) c.d.grav[i] = make_float4(ax, ay, az, phi);
    // interactions = active lanes x listed records - the particles' own leaves
    const uint32_t selfSum = (uint32_t)__popcll(ballot(active && selfListed));
    if (lane == 0) {
        const unsigned long long lanes = (unsigned long long)__popcll(am);
        atomicAdd(&c.cnt.v[0], lanes * (unsigned long long)listed - selfSum);
        if (forced) atomicAdd(&c.cnt.v[1], (unsigned long long)forced);
    }
}

// The launch variants (waves per block, stack depth): (2, 1024), (4, 1024), (4, 512), (8, 512);
// a block's LDS is W wavefront records of 10 KB (8 KB with the shorter stack).
template <int GROUP, int W, int SCAP>
void gravityAs(const DeviceCtx& c, hipStream_t s) {
    const uint32_t grid = (c.slots + (uint32_t)(W * GROUP) - 1u) / (uint32_t)(W * GROUP);
    if (c.prm.theta2 > 0.333333f) kGravity<true, GROUP, W, SCAP><<<grid, W * kWave, 0, s>>>(c);
    else kGravity<false, GROUP, W, SCAP><<<grid, W * kWave, 0, s>>>(c);
}

template <int GROUP>
void gravityGroup(const DeviceCtx& c, hipStream_t s) {
    const uint32_t w = c.tune.gravityWaves, st = c.tune.gravityStack;
    if (w == 2u) gravityAs<GROUP, 2, 1024>(c, s);
    else if (w == 8u) gravityAs<GROUP, 8, 512>(c, s);
    else if (st == 512u) gravityAs<GROUP, 4, 512>(c, s);
    else gravityAs<GROUP, 4, 1024>(c, s);
}

}  // namespace

void launchGravity(const DeviceCtx& c, hipStream_t s) {
    if (c.prm.group == 32u) gravityGroup<32>(c, s);
    else gravityGroup<kWave>(c, s);
}

}  // namespace astraea_hip
