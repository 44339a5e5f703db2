// relax.hip — the per-step corrections of the damped relaxation on the device.
//
// Simulation::relaxLoadedSingleBody (engine/simulation.cpp:585-765) downloads the body after
// every relaxation step, removes its centre of mass and bulk velocity, damps the velocities by
// 0.85, re-anchors u to u_cold(rho) + c_v T_anchor and clears stress and damage, then uploads it
// again.  Here the same corrections run as two kernels per step on every segment at once (each
// segment holds one body): an FP64 reduction of the segment's mass, position and momentum sums,
// and an apply kernel that reads them back on the device.  No host round trip per step.
#include "eos.hpp"
#include "launch.hpp"

namespace astraea_hip {
namespace {

// A block never straddles two segments (segments are 256-aligned).
__device__ __forceinline__ uint32_t blockSegment(const DeviceCtx& c) { return c.seg.slotSeg[blockIdx.x * kBlock]; }

__device__ __forceinline__ void posVel(const DeviceCtx& c, uint32_t i, double (&x)[3], double (&v)[3]) {
    const float4 p = c.cur.pos[i];
    const float4 w = c.cur.vel[i];
    x[0] = p.x; x[1] = p.y; x[2] = p.z;
    v[0] = w.x; v[1] = w.y; v[2] = w.z;
    if (c.fp64) {
        const size_t cap = (size_t)c.stride64;
        for (int a = 0; a < 3; ++a) {
            x[a] = c.cur.p64[a * cap + i];
            v[a] = c.cur.v64[a * cap + i];
        }
    }
}

template <int NQ>
__device__ __forceinline__ void blockAdd(double (&v)[NQ], double* out) {
    __shared__ double sh[kWavesPerBlock][NQ];
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NQ; ++k) v[k] = waveSumD(v[k]);
    if (laneId() == 0)
#pragma unroll
        for (int k = 0; k < NQ; ++k) sh[wave][k] = v[k];
    __syncthreads();
    if (threadIdx.x < NQ) {
        double s = 0.0;
        for (int w = 0; w < kWavesPerBlock; ++w) s += sh[w][threadIdx.x];
        if (s != 0.0) atomicAdd(&out[threadIdx.x], s);
    }
}

__global__ void kZero(double* p, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0.0;
}

// Per segment: sum m, sum m x (3), sum m v (3).
__global__ __launch_bounds__(kBlock) void kSegMoments(DeviceCtx c, double* sums) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t sg = blockSegment(c);
    double v[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) v[k] = 0.0;
    if (i - c.seg.off[sg] < c.seg.n[sg]) {
        double x[3], w[3];
        posVel(c, i, x, w);
        const double m = c.cur.pos[i].w;
        v[0] = m;
        for (int a = 0; a < 3; ++a) {
            v[1 + a] = m * x[a];
            v[4 + a] = m * w[a];
        }
    }
    blockAdd<7>(v, sums + (size_t)sg * 8);
}

// SHIFT: subtract the segment's centre of mass from the positions and its mean velocity from the
// velocities, then scale the velocities by `damping` (simulation.cpp:656-667).  Always: u from
// the anchor temperature of the particle's original id (288 K past the end of the table,
// simulation.cpp:669-674), stress, damage and their rates cleared (686-700).
template <bool SHIFT, bool FP64>
__global__ __launch_bounds__(kBlock) void kRelaxApply(DeviceCtx c, const double* sums, const float* anchor,
                                                      uint32_t nAnchor, double damping) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t sg = blockSegment(c);
    if (i - c.seg.off[sg] >= c.seg.n[sg]) return;
    float4 p = c.cur.pos[i];
    float4 v = c.cur.vel[i];
    if (SHIFT) {
        const double* s = sums + (size_t)sg * 8;
        double com[3] = {0.0, 0.0, 0.0}, vcom[3] = {0.0, 0.0, 0.0};
        if (s[0] > 0.0)
            for (int a = 0; a < 3; ++a) {
                com[a] = s[1 + a] / s[0];
                vcom[a] = s[4 + a] / s[0];
            }
        if (FP64) {
            const size_t cap = (size_t)c.stride64;
            double x[3], w[3];
            for (int a = 0; a < 3; ++a) {
                x[a] = c.cur.p64[a * cap + i] - com[a];
                w[a] = damping * (c.cur.v64[a * cap + i] - vcom[a]);
                c.cur.p64[a * cap + i] = x[a];
                c.cur.v64[a * cap + i] = w[a];
            }
            p.x = (float)x[0]; p.y = (float)x[1]; p.z = (float)x[2];
            v.x = (float)w[0]; v.y = (float)w[1]; v.z = (float)w[2];
        } else {
            p.x = (float)((double)p.x - com[0]);
            p.y = (float)((double)p.y - com[1]);
            p.z = (float)((double)p.z - com[2]);
            v.x = (float)(damping * ((double)v.x - vcom[0]));
            v.y = (float)(damping * ((double)v.y - vcom[1]));
            v.z = (float)(damping * ((double)v.z - vcom[2]));
        }
        c.cur.pos[i] = p;
        c.cur.vel[i] = v;
    }
    const int mat = materialIndex((int)(c.cur.meta[i] & 0xffu));
    const uint32_t oid = c.cur.oid[i];
    const double tAnchor = oid < nAnchor ? (double)anchor[oid] : 288.0;
    float4 hu = c.cur.hu[i];
    hu.z = (float)(coldEnergyPsu(mat, (double)v.w) + material(mat).cv / kUUnit * tAnchor);
    hu.w = 0.0f;
    c.cur.hu[i] = hu;
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    c.cur.str0[i] = z;
    float4 s1 = c.cur.str1[i];
    s1.x = 0.0f; s1.y = 0.0f; s1.z = 0.0f;
    c.cur.str1[i] = s1;
    c.d.dS0[i] = z;
    c.d.dS1[i] = make_float2(0.0f, 0.0f);
    float4 r2 = c.d.rate2[i];
    r2.y = 0.0f;
    c.d.rate2[i] = r2;
}

// Anchor temperature table from the last EOS evaluation: anchor[original id] = T.
__global__ __launch_bounds__(kBlock) void kAnchorFromEos(DeviceCtx c, float* anchor, uint32_t nAnchor) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t sg = blockSegment(c);
    if (i - c.seg.off[sg] >= c.seg.n[sg]) return;
    const uint32_t oid = c.cur.oid[i];
    if (oid < nAnchor) anchor[oid] = c.d.eos[i].z;
}

// Residual motion per segment (simulation.cpp:729-754): sum m, sum m v^2, max |x|^2.
__global__ __launch_bounds__(kBlock) void kRelaxStats(DeviceCtx c, double* out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t sg = blockSegment(c);
    double v[2] = {0.0, 0.0}, r2 = 0.0;
    if (i - c.seg.off[sg] < c.seg.n[sg]) {
        double x[3], w[3];
        posVel(c, i, x, w);
        const double m = c.cur.pos[i].w;
        v[0] = m;
        v[1] = m * (w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    }
    r2 = waveMaxD(r2);
    if (laneId() == 0 && r2 > 0.0)
        atomicMax((unsigned long long*)&out[(size_t)sg * 8 + 2], (unsigned long long)__double_as_longlong(r2));
    blockAdd<2>(v, out + (size_t)sg * 8);
}

}  // namespace

void launchRelaxCorrect(const DeviceCtx& c, hipStream_t s, double* sums, const float* anchor, uint32_t nAnchor,
                        double damping) {
    const uint32_t n = c.seg.count * 8;
    kZero<<<(n + 255) / 256, 256, 0, s>>>(sums, n);
    kSegMoments<<<c.blocks, kBlock, 0, s>>>(c, sums);
    if (c.fp64) kRelaxApply<true, true><<<c.blocks, kBlock, 0, s>>>(c, sums, anchor, nAnchor, damping);
    else kRelaxApply<true, false><<<c.blocks, kBlock, 0, s>>>(c, sums, anchor, nAnchor, damping);
}
void launchReanchor(const DeviceCtx& c, hipStream_t s, const float* anchor, uint32_t nAnchor) {
    kRelaxApply<false, false><<<c.blocks, kBlock, 0, s>>>(c, nullptr, anchor, nAnchor, 1.0);
}
void launchAnchorFromEos(const DeviceCtx& c, hipStream_t s, float* anchor, uint32_t nAnchor) {
    kAnchorFromEos<<<c.blocks, kBlock, 0, s>>>(c, anchor, nAnchor);
}
void launchRelaxStats(const DeviceCtx& c, hipStream_t s, double* out) {
    const uint32_t n = c.seg.count * 8;
    kZero<<<(n + 255) / 256, 256, 0, s>>>(out, n);
    kRelaxStats<<<c.blocks, kBlock, 0, s>>>(c, out);
}

}  // namespace astraea_hip
