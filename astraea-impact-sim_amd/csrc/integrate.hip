// integrate.hip — KDK leapfrog kernels, device-resident timestep, reductions.
//
//   kick     leapfrogKickHalfKernel (engine/sph_gravity.cu:1291-1374): v, rho, u, D^(1/3) (monotone),
//            D, deviatoric stress with trace removal and von Mises / friction yield.  The second
//            kick of a step also evaluates the EOS (the launchEvaluateEos that follows it,
//            simulation.cpp:1732) and the per-particle timestep candidate that
//            launchReduceDtPsu computes at the start of the next step (sph_gravity.cu:1493-1550).
//   drift    leapfrogDriftKernel (sph_gravity.cu:1376-1489): x += dt (v + 0.25 XSPH) over the
//            neighbour table of the last sort; the per-segment scene bounds of the next sort are
//            reduced in the same kernel.
//   dt       the clamp of launchReduceDtPsu (sph_gravity.cu:2136-2161) runs on the device, per
//            segment, so no step waits on a device-to-host copy.
#include "eos.hpp"
#include "launch.hpp"

namespace astraea_hip {
namespace {

__device__ __forceinline__ float dtCandidate(float cfl, float h, float cs, float vx, float vy, float vz, float ax,
                                             float ay, float az, float u, float du, int mat) {
    const float hi = fmaxf(h, 1.0e-6f), csI = fmaxf(cs, 0.05f);
    const float vMag = sqrtf(vx * vx + vy * vy + vz * vz);
    const float dtCfl = cfl * hi / (2.2f * csI + vMag + 1.0e-6f);
    const float aMag = sqrtf(ax * ax + ay * ay + az * az);
    const float dtAcc = 0.25f * sqrtf(hi / (aMag + 1.0e-8f));
    const float uI = fmaxf(u, 0.0f), uScale = fmaxf(uI, u0Ref(mat));
    const float dtTherm = 0.20f * (uI + 0.01f * uScale) / (fabsf(du) + 1.0e-8f);
    return fminf(dtCfl, fminf(dtAcc, dtTherm));
}

// Block reduction of a dt candidate (min) and a radiated power (sum) into the block's segment.
__device__ __forceinline__ void reduceDtRad(const DeviceCtx& c, uint32_t sg, float dtI, double rad) {
    __shared__ float sMin[kWavesPerBlock];
    __shared__ double sRad[kWavesPerBlock];
    const int wave = threadIdx.x >> 6, lane = laneId();
    const float m = waveMinF(dtI);
    const double r = waveSumD(rad);
    if (lane == 0) {
        sMin[wave] = m;
        sRad[wave] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float bm = sMin[0];
        double br = sRad[0];
        for (int w = 1; w < kWavesPerBlock; ++w) {
            bm = fminf(bm, sMin[w]);
            br += sRad[w];
        }
        if (isfinite(bm) && bm > 0.0f) atomicMin(&c.ss.dtRawBits[sg], __float_as_uint(bm));
        atomicAdd(&c.ss.radSum[sg], br);
    }
}

template <bool SECOND, bool CLAMP, bool STRENGTH, bool FP64>
__global__ __launch_bounds__(kBlock) void kKick(DeviceCtx c) {
    if (stepHeld(c)) return;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t sg = c.seg.slotSeg[i];
    const bool active = i - c.seg.off[sg] < c.seg.n[sg];
    float dtI = 1.0e30f;
    double rad = 0.0;
    if (active) {
        const float hdt = 0.5f * c.ss.dt[sg];
        float4 v = c.cur.vel[i];
        const float4 r = c.d.rate[i];
        const float4 r2 = c.d.rate2[i];
        float4 hu = c.cur.hu[i];
        const int mat = (int)(c.cur.meta[i] & 0xffu);
        if (FP64) {
            const size_t cap = (size_t)c.stride64;
            const double h64 = (double)hdt;
            const double vx = c.cur.v64[i] + h64 * (double)r.x;
            const double vy = c.cur.v64[cap + i] + h64 * (double)r.y;
            const double vz = c.cur.v64[2 * cap + i] + h64 * (double)r.z;
            c.cur.v64[i] = vx; c.cur.v64[cap + i] = vy; c.cur.v64[2 * cap + i] = vz;
            v.x = (float)vx; v.y = (float)vy; v.z = (float)vz;
        } else {
            v.x = v.x + hdt * r.x;
            v.y = v.y + hdt * r.y;
            v.z = v.z + hdt * r.z;
        }
        const float rhoNew = fmaxf(v.w + hdt * r.w, 1.0e-9f);
        v.w = rhoNew;
        float uNew = hu.z + hdt * r2.x;
        if (CLAMP) {
            const Material mp = material(mat);
            const double uMin = coldEnergyPsu(mat, (double)rhoNew) + 3.0 * (mp.cv / kUUnit);
            uNew = fmaxf(uNew, (float)uMin);
        } else {
            uNew = fmaxf(uNew, 0.0f);
        }
        float4 s1 = c.cur.str1[i];
        float dc = s1.z + hdt * r2.y;
        dc = fminf(1.0f, fmaxf(s1.z, dc));
        const float D = dc * dc * dc;
        s1.z = dc;
        hu.z = uNew;
        hu.w = D;
        float4 eo = c.d.eos[i];
        if (CLAMP) {
            const EosResult e = evaluateEos(mat, (double)rhoNew, (double)uNew, (double)D, (double)c.prm.gammaHHe);
            eo = make_float4((float)e.p, (float)e.cs, (float)e.T, 0.0f);
            c.d.eos[i] = eo;
            c.d.phase[i] = e.phase;
        }
        if (STRENGTH) {
            const Strength st = strengthOf(mat, c.prm.strengthMask);
            float4 s0;
            if (st.mu <= 0.0f || st.tMelt <= 0.0f || eo.z >= st.tMelt) {
                s0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                s1.x = 0.0f;
                s1.y = 0.0f;
            } else {
                s0 = c.cur.str0[i];
                const float4 d0 = c.d.dS0[i];
                const float2 d1 = c.d.dS1[i];
                float sxx = s0.x + hdt * d0.x, syy = s0.y + hdt * d0.y, szz = s0.z + hdt * d0.z;
                const float sxy = s0.w + hdt * d0.w, sxz = s1.x + hdt * d1.x, syz = s1.y + hdt * d1.y;
                const float tr = (sxx + syy + szz) * (1.0f / 3.0f);
                sxx -= tr; syy -= tr; szz -= tr;
                const float yI = (1.0f - D) * st.y0 * fmaxf(0.0f, 1.0f - eo.z / st.tMelt);
                const float yF = D * fminf(st.y0, 0.6f * fmaxf(0.0f, eo.x));
                const float Y = fmaxf(yI, yF);
                const float j2 = 0.5f * (sxx * sxx + syy * syy + szz * szz + 2.0f * (sxy * sxy + sxz * sxz + syz * syz));
                const float eq = sqrtf(3.0f * j2 + 1.0e-30f);
                const float sc = Y <= 0.0f ? 0.0f : fminf(1.0f, Y / eq);
                s0 = make_float4(sxx * sc, syy * sc, szz * sc, sxy * sc);
                s1.x = sxz * sc;
                s1.y = syz * sc;
            }
            c.cur.str0[i] = s0;
        }
        if (STRENGTH || r2.y != 0.0f) c.cur.str1[i] = s1;
        c.cur.vel[i] = v;
        c.cur.hu[i] = hu;
        if (SECOND) {
            const EosResult e = evaluateEos(mat, (double)rhoNew, (double)uNew, (double)D, (double)c.prm.gammaHHe);
            c.d.eos[i] = make_float4((float)e.p, (float)e.cs, (float)e.T, 0.0f);
            c.d.phase[i] = e.phase;
            const float d = dtCandidate(c.prm.cfl, hu.x, (float)e.cs, v.x, v.y, v.z, r.x, r.y, r.z, uNew, r2.x, mat);
            if (isfinite(d) && d > 0.0f) dtI = d;
            rad = (double)r2.z;
        }
    }
    if (SECOND) reduceDtRad(c, sg, dtI, rad);
}

// dt candidates and radiated power of the current state (after a force evaluation).
__global__ __launch_bounds__(kBlock) void kDtCandidates(DeviceCtx c) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t sg = c.seg.slotSeg[i];
    const bool active = i - c.seg.off[sg] < c.seg.n[sg];
    float dtI = 1.0e30f;
    double rad = 0.0;
    if (active) {
        const float4 v = c.cur.vel[i];
        const float4 r = c.d.rate[i];
        const float4 r2 = c.d.rate2[i];
        const float4 hu = c.cur.hu[i];
        const int mat = (int)(c.cur.meta[i] & 0xffu);
        const float d = dtCandidate(c.prm.cfl, hu.x, c.d.eos[i].y, v.x, v.y, v.z, r.x, r.y, r.z, hu.z, r2.x, mat);
        if (isfinite(d) && d > 0.0f) dtI = d;
        rad = (double)r2.z;
    }
    reduceDtRad(c, sg, dtI, rad);
}

// S: union records staged per round (every value gives the same results; the launcher takes
// the one the context's tuning names).
template <bool FP64, int S>
__global__ __launch_bounds__(kBlock) void kDrift(DeviceCtx c) {
    struct Rec {
        Vec4 p, v;
        float H, invH, wNorm;  // H_j = 2 max(h_j, 1e-6) and its kernel constants, once per record
    };
    __shared__ double sb[kWavesPerBlock][6];
    __shared__ Rec stageBuf[kWavesPerBlock][S];
    if (stepHeld(c)) return;
    const uint32_t b = xcdBlock(blockIdx.x, gridDim.x);
    const uint32_t i = b * kBlock + threadIdx.x;
    const int lane = laneId(), wave = threadIdx.x >> 6;
    const uint32_t sg = c.seg.slotSeg[i];
    const bool active = i - c.seg.off[sg] < c.seg.n[sg];
    float4 pn = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // every lane walks (inactive slots have no entries), so that the wave stages its union together
    const float4 pi = c.cur.pos[i];
    const float4 vi = c.cur.vel[i];
    const float hi = fmaxf(c.cur.hu[i].x, 1.0e-6f), Hi = 2.0f * hi;
    const float rhoI = fmaxf(vi.w, 1.0e-9f);
    float xs = 0.0f, ys = 0.0f, zs = 0.0f;
    const Support sI = supportOf(Hi);
    forEachNeighbour<S>(
        c.nb, i, stageBuf[wave],
        [&](uint32_t j, Rec& r) {
            r.p = loadVec(c.cur.pos, j);
            r.v = loadVec(c.cur.vel, j);
            r.H = c.d.nb[j].x;
        },
        [&](Rec& r) {
            r.H = 2.0f * maxPos(r.H, 1.0e-6f);
            const Support sj = supportOf(r.H);
            r.invH = sj.invH;
            r.wNorm = sj.wNorm;
        },
        [&](const Rec& n) {
            const float4 pj = toFloat4(n.p), vj = toFloat4(n.v);
            const float Hj = n.H;
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const float r2 = dist2f(dx, dy, dz);
            const float Hm = maxPos(Hi, Hj);  // both positive
            if (!(r2 < Hm * Hm)) return;
            const float r = fastSqrt(r2);
            const float wb = 0.5f * (wendlandS(r, sI) + wendlandS(r, Support{n.invH, n.wNorm, 0.0f}));
            const float rb = 0.5f * (rhoI + maxPos(vj.w, 1.0e-9f));
            const float co = (pj.w * fastRcp(rb)) * wb;
            xs += co * (vj.x - vi.x);
            ys += co * (vj.y - vi.y);
            zs += co * (vj.z - vi.z);
        });
    double px = 0.0, py = 0.0, pz = 0.0;
    if (active) {
        const float dt = c.ss.dt[sg];
        if (FP64) {
            const size_t cap = (size_t)c.stride64;
            const double d64 = (double)dt;
            px = c.cur.p64[i] + d64 * (c.cur.v64[i] + 0.25 * (double)xs);
            py = c.cur.p64[cap + i] + d64 * (c.cur.v64[cap + i] + 0.25 * (double)ys);
            pz = c.cur.p64[2 * cap + i] + d64 * (c.cur.v64[2 * cap + i] + 0.25 * (double)zs);
            c.cur.p64[i] = px; c.cur.p64[cap + i] = py; c.cur.p64[2 * cap + i] = pz;
            pn = make_float4((float)px, (float)py, (float)pz, pi.w);
        } else {
            pn = make_float4(pi.x + dt * (vi.x + 0.25f * xs), pi.y + dt * (vi.y + 0.25f * ys),
                             pi.z + dt * (vi.z + 0.25f * zs), pi.w);
            px = pn.x;
            py = pn.y;
            pz = pn.z;
        }
        c.d.posNew[i] = pn;
    }
    // per-block scene bounds of the drifted positions (reduced per segment before the sort): of
    // the FP32 positions, or of the FP64 ones with FP64 accumulators (the Morton box)
    double v[6] = {active ? px : 1.0e30, active ? py : 1.0e30, active ? pz : 1.0e30,
                   active ? px : -1.0e30, active ? py : -1.0e30, active ? pz : -1.0e30};
    for (int a = 0; a < 3; ++a) v[a] = waveMinD(v[a]);
    for (int a = 3; a < 6; ++a) v[a] = waveMaxD(v[a]);
    if (lane == 0)
        for (int a = 0; a < 6; ++a) sb[wave][a] = v[a];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        double r = sb[0][a];
        for (int w = 1; w < kWavesPerBlock; ++w) r = a < 3 ? fmin(r, sb[w][a]) : fmax(r, sb[w][a]);
        c.d.blockBounds[6 * b + a] = r;
    }
}

// Step prologue, one thread per segment: dt from the accumulated minimum (launchReduceDtPsu's
// clamp, sph_gravity.cu:2137-2160), then reset the accumulators of this step.
__global__ void kStepBegin(DeviceCtx c, float maxDt, int resetBounds) {
    const uint32_t sg = blockIdx.x * blockDim.x + threadIdx.x;
    if (sg >= c.seg.count || stepHeld(c)) return;
    const float floorDt = fminf(1.0e-5f, maxDt);
    const float prev = c.ss.dtPrev[sg];
    const float upper = prev > 0.0f ? fminf(prev * 1.15f, maxDt) : maxDt;
    const uint32_t bits = c.ss.dtRawBits[sg];
    float raw = __uint_as_float(bits);
    if (bits == 0x7f7f7f7fu || !isfinite(raw) || raw <= 0.0f) raw = upper;
    const float dt = fminf(upper, fmaxf(floorDt, raw));
    c.ss.dt[sg] = dt;
    c.ss.dtPrev[sg] = dt;
    c.ss.dtRawBits[sg] = 0x7f7f7f7fu;
    c.ss.radSum[sg] = 0.0;
    if (resetBounds) c.ss.living[sg] = 0u;
}

// Step epilogue: radiated energy by the trapezoid rule (simulation.cpp:1741-1744) and time; the
// count of completed steps (the host compares it with the steps it queued).
__global__ void kStepEnd(DeviceCtx c) {
    const uint32_t sg = blockIdx.x * blockDim.x + threadIdx.x;
    if (sg >= c.seg.count || stepHeld(c)) return;
    if (sg == 0u) c.cnt.v[kCntSteps] += 1ull;
    const double dtSec = (double)c.ss.dt[sg] * 1.0e3;
    const double cur = c.ss.radSum[sg];
    c.ss.erad[sg] += 0.5 * (c.ss.radPrev[sg] + cur) * dtSec / 1.0e30;
    c.ss.radPrev[sg] = cur;
    c.ss.simTime[sg] += (double)c.ss.dt[sg];
}

__global__ void kInitIntegrator(DeviceCtx c) {
    const uint32_t sg = blockIdx.x * blockDim.x + threadIdx.x;
    if (sg >= c.seg.count) return;
    c.ss.radPrev[sg] = c.ss.radSum[sg];
    c.ss.erad[sg] = 0.0;
    c.ss.simTime[sg] = 0.0;
    c.ss.dtPrev[sg] = 0.0f;
    c.ss.dt[sg] = 0.0f;
}

// Reset of the per-segment dt minimum and radiated-power sum.
__global__ void kResetDt(DeviceCtx c) {
    const uint32_t sg = blockIdx.x * blockDim.x + threadIdx.x;
    if (sg >= c.seg.count) return;
    c.ss.dtRawBits[sg] = 0x7f7f7f7fu;
    c.ss.radSum[sg] = 0.0;
}

// Reset of the per-segment living-body flag before a sort (sph_gravity.cu:258-271).
__global__ void kResetBounds(DeviceCtx c) {
    const uint32_t sg = blockIdx.x * blockDim.x + threadIdx.x;
    if (sg >= c.seg.count) return;
    c.ss.living[sg] = 0u;
}

// clearStatsKernel (sph_gravity.cu:624-641)
__global__ __launch_bounds__(kBlock) void kClear(DeviceCtx c) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    c.d.grav[i] = z;
    c.d.rate[i] = z;
    c.d.rate2[i] = z;
    c.d.dS0[i] = z;
    c.d.dS1[i] = make_float2(0.0f, 0.0f);
    c.d.p1a[i] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
    c.d.p1b[i] = z;
}

__global__ void kZeroDoubles(double* p, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0.0;
}

// Per-segment FP64 sums for the conservation telemetry (simulation.cpp:985-1087).
__global__ __launch_bounds__(kBlock) void kEnergy(DeviceCtx c, double* out) {
    __shared__ double sh[kWavesPerBlock][14];
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t sg = c.seg.slotSeg[i];
    const bool active = i - c.seg.off[sg] < c.seg.n[sg];
    double v[14];
    for (int k = 0; k < 14; ++k) v[k] = 0.0;
    if (active) {
        const float4 p = c.cur.pos[i];
        const float4 ve = c.cur.vel[i];
        double x = p.x, y = p.y, z = p.z, vx = ve.x, vy = ve.y, vz = ve.z;
        if (c.fp64) {
            const size_t cap = (size_t)c.stride64;
            x = c.cur.p64[i]; y = c.cur.p64[cap + i]; z = c.cur.p64[2 * cap + i];
            vx = c.cur.v64[i]; vy = c.cur.v64[cap + i]; vz = c.cur.v64[2 * cap + i];
        }
        const double m = p.w;
        v[0] = m;
        v[1] = 0.5 * m * (vx * vx + vy * vy + vz * vz);
        v[2] = 0.5 * m * (double)c.d.grav[i].w;
        v[3] = m * (double)c.cur.hu[i].z;
        v[4] = m * vx; v[5] = m * vy; v[6] = m * vz;
        v[7] = m * (y * vz - z * vy);
        v[8] = m * (z * vx - x * vz);
        v[9] = m * (x * vy - y * vx);
        v[10] = m * x; v[11] = m * y; v[12] = m * z;
        v[13] = (double)c.d.rate2[i].z;
    }
    const int wave = threadIdx.x >> 6, lane = laneId();
    for (int k = 0; k < 14; ++k) v[k] = waveSumD(v[k]);
    double mT = 0.0, mP = 0.0;
    if (active) {
        const float4 eo = c.d.eos[i];
        if (isfinite(eo.z)) mT = fmax(0.0, (double)eo.z);
        if (isfinite(eo.x)) mP = fmax(0.0, (double)eo.x);
    }
    mT = waveMaxD(mT);
    mP = waveMaxD(mP);
    __shared__ double sm[kWavesPerBlock][2];
    if (lane == 0) {
        for (int k = 0; k < 14; ++k) sh[wave][k] = v[k];
        sm[wave][0] = mT;
        sm[wave][1] = mP;
    }
    __syncthreads();
    if (threadIdx.x < 14) {
        const int k = threadIdx.x;
        double s = 0.0;
        for (int w = 0; w < kWavesPerBlock; ++w) s += sh[w][k];
        // v[0..9] are ASTRAEA_EN_MASS..ASTRAEA_EN_ANG_Z, v[10..12] the mass-weighted position, v[13] radPowerW
        const int slot = k < 10 ? k : (k < 13 ? ASTRAEA_EN_COM_X + (k - 10) : ASTRAEA_EN_RAD_POWER_W);
        atomicAdd(&out[sg * ASTRAEA_ENERGY_STRIDE + slot], s);
    } else if (threadIdx.x < 16) {
        const int k = threadIdx.x - 14;
        double mx = 0.0;
        for (int w = 0; w < kWavesPerBlock; ++w) mx = fmax(mx, sm[w][k]);
        unsigned long long* dst =
            (unsigned long long*)&out[sg * ASTRAEA_ENERGY_STRIDE + (k == 0 ? ASTRAEA_EN_MAX_TEMP : ASTRAEA_EN_MAX_PRESSURE)];
        atomicMax(dst, (unsigned long long)__double_as_longlong(mx));
    }
}

}  // namespace

void launchKick(const DeviceCtx& c, hipStream_t s, bool second, bool clamp) {
    const dim3 g(c.blocks), t(kBlock);
#define ASTRAEA_KICK(S2, CL, ST, F64) kKick<S2, CL, ST, F64><<<g, t, 0, s>>>(c)
    if (second) {
        if (clamp) {
            if (c.strength) { if (c.fp64) ASTRAEA_KICK(true, true, true, true); else ASTRAEA_KICK(true, true, true, false); }
            else { if (c.fp64) ASTRAEA_KICK(true, true, false, true); else ASTRAEA_KICK(true, true, false, false); }
        } else {
            if (c.strength) { if (c.fp64) ASTRAEA_KICK(true, false, true, true); else ASTRAEA_KICK(true, false, true, false); }
            else { if (c.fp64) ASTRAEA_KICK(true, false, false, true); else ASTRAEA_KICK(true, false, false, false); }
        }
    } else {
        if (clamp) {
            if (c.strength) { if (c.fp64) ASTRAEA_KICK(false, true, true, true); else ASTRAEA_KICK(false, true, true, false); }
            else { if (c.fp64) ASTRAEA_KICK(false, true, false, true); else ASTRAEA_KICK(false, true, false, false); }
        } else {
            if (c.strength) { if (c.fp64) ASTRAEA_KICK(false, false, true, true); else ASTRAEA_KICK(false, false, true, false); }
            else { if (c.fp64) ASTRAEA_KICK(false, false, false, true); else ASTRAEA_KICK(false, false, false, false); }
        }
    }
#undef ASTRAEA_KICK
}

namespace {
template <int S>
void driftAs(const DeviceCtx& c, hipStream_t s) {
    if (c.fp64) kDrift<true, S><<<c.blocks, kBlock, 0, s>>>(c);
    else kDrift<false, S><<<c.blocks, kBlock, 0, s>>>(c);
}
}  // namespace

void launchDrift(const DeviceCtx& c, hipStream_t s) {
    switch (c.tune.driftRound) {
        case 160u: driftAs<160>(c, s); break;
        case 192u: driftAs<192>(c, s); break;
        case 256u: driftAs<256>(c, s); break;
        default: driftAs<128>(c, s); break;
    }
}

void launchStepBegin(const DeviceCtx& c, hipStream_t s, float maxDt) {
    kStepBegin<<<(c.seg.count + 63) / 64, 64, 0, s>>>(c, maxDt, 1);
}
void launchStepEnd(const DeviceCtx& c, hipStream_t s) { kStepEnd<<<(c.seg.count + 63) / 64, 64, 0, s>>>(c); }
void launchDtCandidates(const DeviceCtx& c, hipStream_t s) {
    kResetDt<<<(c.seg.count + 63) / 64, 64, 0, s>>>(c);
    kDtCandidates<<<c.blocks, kBlock, 0, s>>>(c);
}
void launchResetBounds(const DeviceCtx& c, hipStream_t s) { kResetBounds<<<(c.seg.count + 63) / 64, 64, 0, s>>>(c); }
void launchStepBeginNoBounds(const DeviceCtx& c, hipStream_t s, float maxDt) {
    kStepBegin<<<(c.seg.count + 63) / 64, 64, 0, s>>>(c, maxDt, 0);
}
void launchInitIntegrator(const DeviceCtx& c, hipStream_t s) {
    kInitIntegrator<<<(c.seg.count + 63) / 64, 64, 0, s>>>(c);
}
void launchClear(const DeviceCtx& c, hipStream_t s) { kClear<<<c.blocks, kBlock, 0, s>>>(c); }
void launchEnergy(const DeviceCtx& c, hipStream_t s, double* dOut) {
    const uint32_t n = c.seg.count * ASTRAEA_ENERGY_STRIDE;
    kZeroDoubles<<<(n + 255) / 256, 256, 0, s>>>(dOut, n);
    kEnergy<<<c.blocks, kBlock, 0, s>>>(c, dOut);
}

}  // namespace astraea_hip
