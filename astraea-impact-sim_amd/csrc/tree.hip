// tree.hip — scene bounds, Morton keys, permutation + Karras radix tree, bottom-up moments.
//
// Restates launchSortAndBuildBvh (engine/sph_gravity.cu:1996-2061) per segment:
//   bounds   min/max of positions as ordered-uint atomics (sph_gravity.cu:140-149, 273-333);
//   Morton   q = clamp((x - min) * (2097151 / span), 0, 2097151) truncated, 21-bit spread,
//            key = X << 2 | Y << 1 | Z (sph_gravity.cu:151-171, 335-361);
//   permute  gather of the persistent state by the sorted slot (sph_gravity.cu:363-430), fused
//            with the Karras build (sph_gravity.cu:432-529), which needs only the sorted keys;
//   moments  leaf-to-root climb with an arrival counter per node (sph_gravity.cu:531-620), in
//            LDS per chunk of leaves, then over the nodes spanning chunks in a second kernel
//            with agent-scope counters, so the second arriver sees the first one's node record
//            across the 8 XCD-private L2 caches.
// All FP32 arithmetic that decides integer results (quantisation, combine of moments) is
// written with explicit operation order; the oracle repeats it bit for bit.
#include "launch.hpp"

namespace astraea_hip {
namespace {

__device__ __forceinline__ uint64_t spread21(uint64_t v) {
    v &= 0x1fffffULL;
    v = (v | (v << 32)) & 0x001f00000000ffffULL;
    v = (v | (v << 16)) & 0x001f0000ff0000ffULL;
    v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}

__device__ __forceinline__ uint32_t quantise(float x, float mn, float inv) {
#pragma clang fp contract(off)
    const float f = fminf(fmaxf((x - mn) * inv, 0.0f), 2097151.0f);
    return (uint32_t)f;
}

__device__ __forceinline__ float invSpan(float mn, float mx) {
#pragma clang fp contract(off)
    const float span = mx - mn;
    return span > 1.0e-6f ? __fdiv_rn(2097151.0f, span) : 0.0f;
}

// The same quantisation in FP64, from the FP64 positions and an FP64 box (the mixed form with
// FP64 leapfrog accumulators): 21 bits per axis whatever the box's distance from the origin.
__device__ __forceinline__ uint32_t quantise64(double x, double mn, double inv) {
#pragma clang fp contract(off)
    const double f = fmin(fmax((x - mn) * inv, 0.0), 2097151.0);
    return (uint32_t)f;
}

__device__ __forceinline__ double invSpan64(double mn, double mx) {
#pragma clang fp contract(off)
    const double span = mx - mn;
    return span > 1.0e-6 ? 2097151.0 / span : 0.0;
}

__device__ __forceinline__ int karrasDelta(const uint64_t* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = k[i], b = k[j];
    if (a != b) return __clzll((long long)(a ^ b));
    return 64 + __clz((int)((uint32_t)i ^ (uint32_t)j));
}

// Block min/max of the positions into the block's partial-bounds record.
__device__ __forceinline__ void blockBounds(double x, double y, double z, bool active, double* out6) {
    __shared__ double sb[kWavesPerBlock][6];
    const int wave = threadIdx.x >> 6, lane = laneId();
    double v[6] = {active ? x : 1.0e30, active ? y : 1.0e30, active ? z : 1.0e30,
                   active ? x : -1.0e30, active ? y : -1.0e30, active ? z : -1.0e30};
    for (int a = 0; a < 3; ++a) v[a] = waveMinD(v[a]);
    for (int a = 3; a < 6; ++a) v[a] = waveMaxD(v[a]);
    if (lane == 0)
        for (int a = 0; a < 6; ++a) sb[wave][a] = v[a];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        double r = sb[0][a];
        for (int w = 1; w < kWavesPerBlock; ++w) r = a < 3 ? fmin(r, sb[w][a]) : fmax(r, sb[w][a]);
        out6[a] = r;
    }
}

// posNew = current positions (the sort input when no drift precedes it) + bounds.
__global__ __launch_bounds__(kBlock) void kPrepareSort(DeviceCtx c) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t sg = c.seg.slotSeg[i];
    const bool active = i - c.seg.off[sg] < c.seg.n[sg];
    const float4 p = c.cur.pos[i];  // in the FP64 mode pos always holds the rounded FP64 position
    if (active) c.d.posNew[i] = p;
    if (c.fp64) {
        const size_t cap = (size_t)c.stride64;
        blockBounds(c.cur.p64[i], c.cur.p64[cap + i], c.cur.p64[2 * cap + i], active, c.d.blockBounds + 6 * blockIdx.x);
    } else {
        blockBounds(p.x, p.y, p.z, active, c.d.blockBounds + 6 * blockIdx.x);
    }
}

__global__ __launch_bounds__(kBlock) void kMorton(DeviceCtx c) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t sg = c.seg.slotSeg[i];
    const bool active = i - c.seg.off[sg] < c.seg.n[sg];
    uint64_t key = kPadKey;
    if (active) {
        const double* bo = c.ss.bounds + 6 * sg;
        if (c.fp64) {
            // the mixed form: FP64 box and FP64 positions (the leapfrog's masters)
            const size_t cap = (size_t)c.stride64;
            const double x = c.cur.p64[i], y = c.cur.p64[cap + i], z = c.cur.p64[2 * cap + i];
            key = (spread21(quantise64(x, bo[0], invSpan64(bo[0], bo[3]))) << 2) |
                  (spread21(quantise64(y, bo[1], invSpan64(bo[1], bo[4]))) << 1) |
                  spread21(quantise64(z, bo[2], invSpan64(bo[2], bo[5])));
        } else {
            const float mnx = (float)bo[0], mny = (float)bo[1], mnz = (float)bo[2];  // FP32 values, exactly
            const float mxx = (float)bo[3], mxy = (float)bo[4], mxz = (float)bo[5];
            const float4 p = c.d.posNew[i];
            key = (spread21(quantise(p.x, mnx, invSpan(mnx, mxx))) << 2) |
                  (spread21(quantise(p.y, mny, invSpan(mny, mxy))) << 1) | spread21(quantise(p.z, mnz, invSpan(mnz, mxz)));
        }
    }
    c.sb.keyA[i] = key;
    c.sb.valA[i] = i;
}

// Gather the persistent state into the sorted order and build the Karras tree of each segment.
// While the step sequence is held (stepHeld) the state is copied in place and the tree is kept:
// the host swaps the two sets after this kernel either way.
__global__ __launch_bounds__(kBlock) void kPermuteKarras(DeviceCtx c) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t sg = c.seg.slotSeg[i];
    const uint32_t off = c.seg.off[sg], n = c.seg.n[sg];
    const uint32_t local = i - off;
    if (local >= n) return;
    const bool hold = stepHeld(c);
    const uint32_t src = hold ? i : c.sb.valA[i];
    const size_t cap = (size_t)c.stride64;
    float4 p = hold ? c.cur.pos[i] : c.d.posNew[src];
    if (c.fp64) {
        const double x = c.cur.p64[src], y = c.cur.p64[cap + src], z = c.cur.p64[2 * cap + src];
        c.alt.p64[i] = x; c.alt.p64[cap + i] = y; c.alt.p64[2 * cap + i] = z;
        c.alt.v64[i] = c.cur.v64[src]; c.alt.v64[cap + i] = c.cur.v64[cap + src]; c.alt.v64[2 * cap + i] = c.cur.v64[2 * cap + src];
        p.x = (float)x; p.y = (float)y; p.z = (float)z;
    }
    c.alt.pos[i] = p;
    c.alt.vel[i] = c.cur.vel[src];
    const float4 hu = c.cur.hu[src];
    c.alt.hu[i] = hu;
    c.alt.str0[i] = c.cur.str0[src];
    c.alt.str1[i] = c.cur.str1[src];
    const uint32_t meta = c.cur.meta[src];
    c.alt.meta[i] = meta;
    c.alt.oid[i] = c.cur.oid[src];
    if (hold) return;
    if (((meta >> 16) & 0xffu) != 0u && hu.w < 0.8f) atomicOr(&c.ss.living[sg], 1u);

    if (n == 1u) {
        c.t.parentL[i] = -1;
        return;
    }
    if (local + 1u < n) {
        const uint64_t* k = c.sb.keyA + off;
        const int li = (int)local, ni = (int)n;
        const int d = (karrasDelta(k, ni, li, li + 1) - karrasDelta(k, ni, li, li - 1)) >= 0 ? 1 : -1;
        const int dmin = karrasDelta(k, ni, li, li - d);
        int lmax = 2;
        while (karrasDelta(k, ni, li, li + lmax * d) > dmin) lmax <<= 1;
        int l = 0;
        for (int t = lmax >> 1; t >= 1; t >>= 1)
            if (karrasDelta(k, ni, li, li + (l + t) * d) > dmin) l += t;
        const int j = li + l * d;
        const int dnode = karrasDelta(k, ni, li, j);
        int s = 0, t = l;
        do {
            t = (t + 1) >> 1;
            if (karrasDelta(k, ni, li, li + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        const int gamma = li + s * d + (d < 0 ? d : 0);
        const int first = li < j ? li : j, last = li > j ? li : j;
        const int o = (int)off;
        const bool lLeaf = first == gamma, rLeaf = last == gamma + 1;
        c.t.child[i] = make_int2(lLeaf ? ~(o + gamma) : o + gamma, rLeaf ? ~(o + gamma + 1) : o + gamma + 1);
        if (lLeaf) c.t.parentL[o + gamma] = (int)i; else c.t.parentI[o + gamma] = (int)i;
        if (rLeaf) c.t.parentL[o + gamma + 1] = (int)i; else c.t.parentI[o + gamma + 1] = (int)i;
        c.t.visited[i] = 0u;
        c.t.other[i] = o + j;
        if (local == 0u) c.t.parentI[i] = -1;
    }
}

constexpr int kChunk = 512;  // leaf slots per block of the moments climb

struct NodeRec {
    float4 cm, mn, mx;
};

// Node records written and read during the climb use agent-scope relaxed atomics: on gfx950
// they are global stores/loads with the sc1 coherence bit, visible across the XCD-private L2
// caches without the whole-L2 write-back/invalidate that an acq_rel fence emits.
__device__ __forceinline__ void storeAgent(float4* dst, float4 v) {
    float* d = reinterpret_cast<float*>(dst);
    __hip_atomic_store(d + 0, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(d + 1, v.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(d + 2, v.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(d + 3, v.w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float4 loadAgent(const float4* src) {
    float* s = const_cast<float*>(reinterpret_cast<const float*>(src));
    float4 v;
    v.x = __hip_atomic_load(s + 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    v.y = __hip_atomic_load(s + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    v.z = __hip_atomic_load(s + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    v.w = __hip_atomic_load(s + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return v;
}

__device__ __forceinline__ NodeRec loadChild(const DeviceCtx& c, int ch) {
    NodeRec r;
    if (ch < 0) {
        const int j = ~ch;
        const float4 p = c.cur.pos[j];
        const float4 hu = c.cur.hu[j];
        r.cm = p;
        r.mn = make_float4(p.x, p.y, p.z, 0.0f);
        r.mx = make_float4(p.x, p.y, p.z, supportBound(hu.x, hu.y));
    } else {
        r.cm = loadAgent(&c.t.cm[ch]);
        r.mn = loadAgent(&c.t.bmin[ch]);
        r.mx = loadAgent(&c.t.bmax[ch]);
    }
    return r;
}

// Combine of two children (the FP32 replica the oracle reproduces bit for bit).
__device__ __forceinline__ NodeRec combine(const NodeRec& L, const NodeRec& R) {
#pragma clang fp contract(off)
    NodeRec o;
    o.mn.x = fminf(L.mn.x, R.mn.x); o.mn.y = fminf(L.mn.y, R.mn.y); o.mn.z = fminf(L.mn.z, R.mn.z);
    o.mx.x = fmaxf(L.mx.x, R.mx.x); o.mx.y = fmaxf(L.mx.y, R.mx.y); o.mx.z = fmaxf(L.mx.z, R.mx.z);
    o.mx.w = fmaxf(L.mx.w, R.mx.w);
    const float mS = L.cm.w + R.cm.w;
    if (mS > 0.0f) {
        const float invM = __fdiv_rn(1.0f, mS);
        o.cm.x = __builtin_fmaf(L.cm.w, L.cm.x, R.cm.w * R.cm.x) * invM;
        o.cm.y = __builtin_fmaf(L.cm.w, L.cm.y, R.cm.w * R.cm.y) * invM;
        o.cm.z = __builtin_fmaf(L.cm.w, L.cm.z, R.cm.w * R.cm.z) * invM;
    } else {
        o.cm.x = 0.5f * (L.cm.x + R.cm.x);
        o.cm.y = 0.5f * (L.cm.y + R.cm.y);
        o.cm.z = 0.5f * (L.cm.z + R.cm.z);
    }
    o.cm.w = mS;
    o.mn.w = fmaxf(o.mx.x - o.mn.x, fmaxf(o.mx.y - o.mn.y, o.mx.z - o.mn.z));
    return o;
}

__device__ __forceinline__ bool spansChunks(const DeviceCtx& c, int p) {
    const int o = c.t.other[p];
    const int lo = p < o ? p : o, hi = p < o ? o : p;
    return lo / kChunk != hi / kChunk;
}

// Leaf-to-root climb over the permuted state (the host has swapped the sets: `cur` is sorted), in
// two kernels.  The first thread to reach a node stops; the second combines both children.
//
// kMoments: a block takes a chunk of kChunk consecutive leaf slots.  A node whose leaves all lie in
// the chunk (its index is one end of its range, so it is in the chunk too) is reached only by the
// block's own threads.  The block first loads what the climb reads into LDS with coalesced loads:
// each slot's leaf (position, support bound) and internal node (children, parent, far end of its
// range), so that a level of the climb costs LDS accesses only, not three dependent global loads;
// arrival counters and node records live in LDS too.  A thread that reaches a node spanning two
// chunks (a few percent of the nodes, all near the top of the tree) counts its arrival there and
// stops, so no block waits for a climb through the top of the tree.
//
// kMomentsTop: the nodes spanning chunks, climbed by one thread per node whose children were both
// finished by kMoments; its arrivals add to the counts kMoments left, so a node's second arrival
// (of two children, however they were finished) continues.  Records are written and read with
// agent-scope relaxed atomics (global stores/loads with the sc1 coherence bit, visible across the
// XCD-private L2 caches without a whole-L2 write-back), and a thread's record stores completed
// (s_waitcnt) before its arrival increment.
__global__ __launch_bounds__(kChunk) void kMoments(DeviceCtx c) {
    __shared__ uint32_t arrive[kChunk];
    __shared__ NodeRec rec[kChunk];
    __shared__ int2 kid[kChunk];
    __shared__ int up[kChunk];
    __shared__ int far[kChunk];
    __shared__ float4 leafP[kChunk];
    __shared__ float leafH[kChunk];
    if (stepHeld(c)) return;
    const int base = (int)(blockIdx.x * kChunk);
    const int t = (int)threadIdx.x;
    const uint32_t i = (uint32_t)base + (uint32_t)t;
    arrive[t] = 0u;
    int p = -1;
    if (i < c.slots) {
        const uint32_t sg = c.seg.slotSeg[i];
        const uint32_t n = c.seg.n[sg], local = i - c.seg.off[sg];
        if (n >= 2u && local < n) {
            leafP[t] = c.cur.pos[i];
            const float4 hu = c.cur.hu[i];
            leafH[t] = supportBound(hu.x, hu.y);
            p = c.t.parentL[i];
            if (local + 1u < n) {  // slot i also holds internal node i
                kid[t] = c.t.child[i];
                up[t] = c.t.parentI[i];
                far[t] = c.t.other[i];
            }
        }
    }
    __syncthreads();
    const auto leafRecAt = [&](int u) {
        const float4 q = leafP[u];
        return NodeRec{q, make_float4(q.x, q.y, q.z, 0.0f), make_float4(q.x, q.y, q.z, leafH[u])};
    };
    while (p >= 0) {
        // node p lies in the chunk when both ends of its range do (p is one end, far[] the other)
        const int q = p - base;
        bool in = (uint32_t)q < (uint32_t)kChunk;
        if (in) in = (uint32_t)(far[q] - base) < (uint32_t)kChunk;
        if (!in) {
            __hip_atomic_fetch_add(&c.t.visited[p], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        const uint32_t prior = __hip_atomic_fetch_add(&arrive[q], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (prior == 0u) return;
        const int2 ch = kid[q];
        const NodeRec L = ch.x < 0 ? leafRecAt(~ch.x - base) : rec[ch.x - base];
        const NodeRec R = ch.y < 0 ? leafRecAt(~ch.y - base) : rec[ch.y - base];
        const NodeRec out = combine(L, R);
        rec[q] = out;
        c.t.cm[p] = out.cm;
        c.t.bmin[p] = out.mn;
        c.t.bmax[p] = out.mx;
        p = up[q];
    }
}

__global__ __launch_bounds__(kBlock) void kMomentsTop(DeviceCtx c) {
    if (stepHeld(c)) return;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t sg = c.seg.slotSeg[i];
    const uint32_t n = c.seg.n[sg];
    if (n < 2u || i - c.seg.off[sg] >= n - 1u) return;  // internal nodes only
    int p = (int)i;
    if (!spansChunks(c, p)) return;
    int2 ch = c.t.child[p];
    // a child spanning chunks is finished here, by the climb that reaches p through it
    if ((ch.x >= 0 && spansChunks(c, ch.x)) || (ch.y >= 0 && spansChunks(c, ch.y))) return;
    NodeRec out = combine(loadChild(c, ch.x), loadChild(c, ch.y));
    int pp = c.t.parentI[p];
    // a level costs three round trips to memory: the record's stores (with the parent's topology
    // loaded meanwhile), the arrival, and the sibling's record (the node's own is in registers)
    for (;;) {
        storeAgent(&c.t.cm[p], out.cm);
        storeAgent(&c.t.bmin[p], out.mn);
        storeAgent(&c.t.bmax[p], out.mx);
        if (pp < 0) return;
        const int2 up = c.t.child[pp];
        const int upp = c.t.parentI[pp];
        __builtin_amdgcn_s_waitcnt(0);
        const uint32_t prior = __hip_atomic_fetch_add(&c.t.visited[pp], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (prior == 0u) return;  // the other child is not finished yet: its climb continues
        // p is an internal node (>= 0): a leaf child is stored as ~slot (< 0)
        const bool left = up.x == p;
        const NodeRec sib = loadChild(c, left ? up.y : up.x);
        out = left ? combine(out, sib) : combine(sib, out);
        p = pp;
        pp = upp;
    }
}

// Scene bounds of each segment from its blocks' partial records (min/max is exact in any order).
__global__ __launch_bounds__(kBlock) void kBoundsReduce(DeviceCtx c) {
    __shared__ double sb[kWavesPerBlock][6];
    const uint32_t sg = blockIdx.x;
    const uint32_t b0 = c.seg.off[sg] / kBlock, b1 = c.seg.off[sg + 1] / kBlock;
    double v[6] = {1.0e30, 1.0e30, 1.0e30, -1.0e30, -1.0e30, -1.0e30};
    for (uint32_t b = b0 + threadIdx.x; b < b1; b += kBlock) {
        const double* p = c.d.blockBounds + 6 * b;
        for (int a = 0; a < 3; ++a) {
            v[a] = fmin(v[a], p[a]);
            v[3 + a] = fmax(v[3 + a], p[3 + a]);
        }
    }
    for (int a = 0; a < 3; ++a) v[a] = waveMinD(v[a]);
    for (int a = 3; a < 6; ++a) v[a] = waveMaxD(v[a]);
    const int wave = threadIdx.x >> 6;
    if (laneId() == 0)
        for (int a = 0; a < 6; ++a) sb[wave][a] = v[a];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        double r = sb[0][a];
        for (int w = 1; w < kWavesPerBlock; ++w) r = a < 3 ? fmin(r, sb[w][a]) : fmax(r, sb[w][a]);
        c.ss.bounds[6 * sg + a] = r;
    }
}

}  // namespace

void launchBoundsReduce(const DeviceCtx& c, hipStream_t s) { kBoundsReduce<<<c.seg.count, kBlock, 0, s>>>(c); }

void launchPrepareSort(const DeviceCtx& c, hipStream_t s) { kPrepareSort<<<c.blocks, kBlock, 0, s>>>(c); }
void launchMorton(const DeviceCtx& c, hipStream_t s) { kMorton<<<c.blocks, kBlock, 0, s>>>(c); }
void launchPermuteKarras(const DeviceCtx& c, hipStream_t s) { kPermuteKarras<<<c.blocks, kBlock, 0, s>>>(c); }
void launchMoments(const DeviceCtx& c, hipStream_t s) {
    kMoments<<<(c.slots + kChunk - 1) / kChunk, kChunk, 0, s>>>(c);
    kMomentsTop<<<c.blocks, kBlock, 0, s>>>(c);
}

}  // namespace astraea_hip
