// telemetry.hip — the largest-remnant classification of the telemetry on the device.
//
// Restates the particle loops of Simulation::computeTelemetryFromDevice
// (engine/simulation.cpp:1000-1227), which the reference runs on the host over every particle
// each step, as per-segment FP64 reductions:
//   * body sums: mass, mass-weighted position and velocity of body 0 and of the other bodies
//     (lines 1000-1013), and the seed of the remnant search, the particle of lowest phi
//     (lines 1036-1039; the first in input order, i.e. lowest original id, on ties);
//   * one remnant iteration: the mass, centre, velocity and sum m r^2 of the particles bound to
//     the current remnant guess, e = v^2/2 - G M / max(r, r_soft) < 0 (lines 1131-1178);
//   * the final split into bound planet / disk / ejecta, the bound impactor mass and the largest
//     damage (lines 1180-1227), and the spin of body 0 within 1.5 target radii (lines 1094-1112).
// The host runs the five remnant iterations (the reference's loop and its stop rule) between
// the reductions.  Block partials are combined with FP64 atomics, one per block and quantity.
#include "launch.hpp"

namespace astraea_hip {
namespace {

constexpr double kGPsu = 66.7430;  // engine/units.hpp:26

struct PV {
    double x, y, z, vx, vy, vz, m;
};

__device__ __forceinline__ PV stateOf(const DeviceCtx& c, uint32_t i) {
    const float4 p = c.cur.pos[i];
    const float4 v = c.cur.vel[i];
    PV s{p.x, p.y, p.z, v.x, v.y, v.z, p.w};
    if (c.fp64) {
        const size_t cap = (size_t)c.stride64;
        s.x = c.cur.p64[i]; s.y = c.cur.p64[cap + i]; s.z = c.cur.p64[2 * cap + i];
        s.vx = c.cur.v64[i]; s.vy = c.cur.v64[cap + i]; s.vz = c.cur.v64[2 * cap + i];
    }
    return s;
}

// Sums NQ doubles over the block and adds them to out[sg * stride + slot0 + k].
template <int NQ>
__device__ __forceinline__ void blockAdd(double (&v)[NQ], double* out, uint32_t sg, int stride, int slot0) {
    __shared__ double sh[kWavesPerBlock][NQ];
    const int wave = threadIdx.x >> 6, lane = laneId();
#pragma unroll
    for (int k = 0; k < NQ; ++k) v[k] = waveSumD(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NQ; ++k) sh[wave][k] = v[k];
    __syncthreads();
    if (threadIdx.x < NQ) {
        double s = 0.0;
        for (int w = 0; w < kWavesPerBlock; ++w) s += sh[w][threadIdx.x];
        if (s != 0.0) atomicAdd(&out[(size_t)sg * stride + slot0 + threadIdx.x], s);
    }
}

// A block never straddles two segments (segments are 256-aligned), so its segment is uniform.
__device__ __forceinline__ uint32_t blockSegment(const DeviceCtx& c) { return c.seg.slotSeg[blockIdx.x * kBlock]; }

__global__ __launch_bounds__(kBlock) void kBodySums(DeviceCtx c, double* out, unsigned long long* seed) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t sg = blockSegment(c);
    const uint32_t off = c.seg.off[sg];
    const bool active = i - off < c.seg.n[sg];
    double v[14];
#pragma unroll
    for (int k = 0; k < 14; ++k) v[k] = 0.0;
    unsigned long long key = ~0ull;
    if (active) {
        const PV s = stateOf(c, i);
        const int b = ((c.cur.meta[i] >> 8) & 0xffu) == 0u ? 0 : 7;
        v[b + 0] = s.m;
        v[b + 1] = s.m * s.x; v[b + 2] = s.m * s.y; v[b + 3] = s.m * s.z;
        v[b + 4] = s.m * s.vx; v[b + 5] = s.m * s.vy; v[b + 6] = s.m * s.vz;
        // lowest finite phi, then lowest original id: ordered float bits above the id.  A
        // non-finite phi ranks after every finite one, so a segment without any seeds at its
        // first particle, as the reference's search does.
        const float phi = c.d.grav[i].w;
        const unsigned long long hi = isfinite(phi) ? floatToOrdered(phi) : 0xffffffffu;
        key = (hi << 32) | (unsigned long long)c.cur.oid[i];
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(key, o, 64);
        key = w < key ? w : key;
    }
    if (laneId() == 0 && key != ~0ull) atomicMin(&seed[sg], key);
    blockAdd<14>(v, out, sg, ASTRAEA_CLASS_STRIDE, ASTRAEA_CL_BODY0_MASS);
}

// The seed particle (the one whose original id is in the low half of its segment's key)
// writes its position and velocity.  Original ids are unique, so one thread writes.
__global__ __launch_bounds__(kBlock) void kSeedState(DeviceCtx c, const unsigned long long* seed, double* out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t sg = blockSegment(c);
    const unsigned long long key = seed[sg];
    if (key == ~0ull || i - c.seg.off[sg] >= c.seg.n[sg] || c.cur.oid[i] != (uint32_t)key) return;
    const PV s = stateOf(c, i);
    double* o = out + (size_t)sg * ASTRAEA_CLASS_STRIDE + ASTRAEA_CL_SEED_X;
    o[0] = s.x; o[1] = s.y; o[2] = s.z; o[3] = s.vx; o[4] = s.vy; o[5] = s.vz;
}

__global__ void kFill(double* p, uint32_t n, unsigned long long* q, uint32_t nq) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0.0;
    if (k < nq) q[k] = ~0ull;
}

// One remnant iteration.  rem[sg*8 + ...] = (mass, x, y, z, vx, vy, vz, radius); a segment with a
// negative mass has stopped iterating and is skipped.
__global__ __launch_bounds__(kBlock) void kRemnant(DeviceCtx c, const double* rem, double* out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t sg = blockSegment(c);
    const double* r = rem + (size_t)sg * 8;
    if (!(r[0] >= 0.0)) return;  // uniform over the block
    const bool active = i - c.seg.off[sg] < c.seg.n[sg];
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = 0.0;
    if (active) {
        const PV s = stateOf(c, i);
        const double dx = s.x - r[1], dy = s.y - r[2], dz = s.z - r[3];
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double rr = sqrt(r2);
        const double dvx = s.vx - r[4], dvy = s.vy - r[5], dvz = s.vz - r[6];
        const double v2 = dvx * dvx + dvy * dvy + dvz * dvz;
        const double rSoft = fmax(0.5 * r[7], 0.1);
        const double e = 0.5 * v2 - kGPsu * r[0] / fmax(rr, rSoft);
        if (e < 0.0) {
            v[0] = s.m;
            v[1] = s.m * s.x; v[2] = s.m * s.y; v[3] = s.m * s.z;
            v[4] = s.m * s.vx; v[5] = s.m * s.vy; v[6] = s.m * s.vz;
            v[7] = s.m * r2;
        }
    }
    blockAdd<8>(v, out, sg, 8, 0);
}

// Final split.  rem as in kRemnant (mass >= 0 always); body 0's centre and velocity and the
// target radius are in body[sg*8 + ...] = (x0, y0, z0, vx0, vy0, vz0, target radius, 0).
__global__ __launch_bounds__(kBlock) void kClassify(DeviceCtx c, const double* rem, const double* body, double* out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t sg = blockSegment(c);
    const double* r = rem + (size_t)sg * 8;
    const double* b0 = body + (size_t)sg * 8;
    const bool active = i - c.seg.off[sg] < c.seg.n[sg];
    double v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = 0.0;
    double dmg = 0.0;
    if (active) {
        const PV s = stateOf(c, i);
        dmg = fmax(0.0, (double)c.cur.hu[i].w);
        const bool impactor = ((c.cur.meta[i] >> 8) & 0xffu) != 0u;
        const double rSoft = fmax(0.5 * r[7], 0.1);
        const double mu = kGPsu * r[0];
        const double dx = s.x - r[1], dy = s.y - r[2], dz = s.z - r[3];
        const double rr = sqrt(dx * dx + dy * dy + dz * dz);
        const double dvx = s.vx - r[4], dvy = s.vy - r[5], dvz = s.vz - r[6];
        const double v2 = dvx * dvx + dvy * dvy + dvz * dvz;
        const double e = 0.5 * v2 - mu / fmax(rr, rSoft);
        if (e >= 0.0) {
            v[2] = s.m;  // ejecta
        } else {
            if (impactor) v[3] = s.m;
            bool disk = false;
            if (rr > r[7] && mu > 0.0) {
                const double jx = dy * dvz - dz * dvy, jy = dz * dvx - dx * dvz, jz = dx * dvy - dy * dvx;
                const double j2 = jx * jx + jy * jy + jz * jz;
                const double ecc = sqrt(fmax(0.0, 1.0 + (2.0 * e * j2) / (mu * mu)));
                disk = j2 / (mu * (1.0 + ecc)) > r[7];
            }
            if (disk) v[1] = s.m; else v[0] = s.m;
        }
        // spin of body 0 about its centre, within 1.5 target radii (simulation.cpp:1094-1112)
        if (!impactor) {
            const double ex = s.x - b0[0], ey = s.y - b0[1], ez = s.z - b0[2];
            const double lim = 1.5 * b0[6];
            if (ex * ex + ey * ey + ez * ez <= lim * lim) {
                v[4] = s.m * (ez * (s.vx - b0[3]) - ex * (s.vz - b0[5]));
                v[5] = s.m * (ex * ex + ez * ez);
            }
        }
    }
    dmg = waveMaxD(dmg);
    if (laneId() == 0 && dmg > 0.0)
        atomicMax((unsigned long long*)&out[(size_t)sg * ASTRAEA_CLASS_STRIDE + ASTRAEA_CL_MAX_DAMAGE],
                  (unsigned long long)__double_as_longlong(dmg));
    blockAdd<6>(v, out, sg, ASTRAEA_CLASS_STRIDE, ASTRAEA_CL_BOUND_PLANET);
}

}  // namespace

void launchClassBodies(const DeviceCtx& c, hipStream_t s, double* out, unsigned long long* seed) {
    const uint32_t n = c.seg.count * ASTRAEA_CLASS_STRIDE;
    kFill<<<(n + 255) / 256, 256, 0, s>>>(out, n, seed, c.seg.count);
    kBodySums<<<c.blocks, kBlock, 0, s>>>(c, out, seed);
    kSeedState<<<c.blocks, kBlock, 0, s>>>(c, seed, out);
}
void launchRemnant(const DeviceCtx& c, hipStream_t s, const double* rem, double* acc) {
    const uint32_t n = c.seg.count * 8;
    kFill<<<(n + 255) / 256, 256, 0, s>>>(acc, n, nullptr, 0);
    kRemnant<<<c.blocks, kBlock, 0, s>>>(c, rem, acc);
}
void launchClassify(const DeviceCtx& c, hipStream_t s, const double* rem, const double* body, double* out) {
    kClassify<<<c.blocks, kBlock, 0, s>>>(c, rem, body, out);
}

}  // namespace astraea_hip
