// sph.hip — SPH pass 1 (+ fused EOS) and pass 2 over the neighbour table.
//
// Pass 1 restates sphPass1Kernel (engine/sph_gravity.cu:645-788): over r < H_i (self
// excluded, count starting at 1) it gathers Shepard C, the surface normal and the velocity
// gradient, sets f_surf, div v, curl v and the new h.  Its epilogue evaluates the EOS of the
// same particle (evaluateEosKernel, sph_gravity.cu:792-806: its inputs rho, u, D are not
// changed by pass 1) and stores the per-particle terms pass 2 reads for every neighbour:
// h', P_eff/rho^2, max(c_s, 0.05), the Balsara switch (sph_gravity.cu:222-228, 965-969) and,
// with strength, S/rho^2.
//
// Pass 2 restates sphPass2Kernel (engine/sph_gravity.cu:971-1287): over r < max(H_i, H_j),
// delta-SPH continuity with same-material diffusion, Monaghan viscosity with the Balsara
// switch, the tensile correction, pressure and deviatoric-stress forces, du/dt with
// radiative cooling (FP64, 1219-1235), the Jaumann stress rate and Grady-Kipp damage rate.
#include "eos.hpp"
#include "launch.hpp"

namespace astraea_hip {
namespace {

__device__ __forceinline__ float balsaraSwitch(float divV, float cx, float cy, float cz, float cs, float h) {
    const float ad = fabsf(divV), ac = sqrtf(cx * cx + cy * cy + cz * cz);
    const float den = ad + ac + 1.0e-4f * fmaxf(cs, 0.05f) / fmaxf(h, 1.0e-6f);
    return ad / fmaxf(den, 1.0e-30f);
}

__device__ __forceinline__ float effectivePressure(float p, uint8_t phase) {
    return phase != kSolid ? fmaxf(p, 0.0f) : p;
}

// Largest principal stress of -P I + S for traceless S (trigonometric root of the
// characteristic cubic), sph_gravity.cu:230-254.
__device__ __forceinline__ float maxPrincipalStress(float p, float sxx, float syy, float szz, float sxy, float sxz,
                                                    float syz) {
    const float p2 = sxx * sxx + syy * syy + szz * szz + 2.0f * (sxy * sxy + sxz * sxz + syz * syz);
    if (p2 <= 1.0e-24f) return -p;
    const float q = sqrtf(p2 * (1.0f / 6.0f)), iq = 1.0f / q;
    const float bxx = sxx * iq, byy = syy * iq, bzz = szz * iq, bxy = sxy * iq, bxz = sxz * iq, byz = syz * iq;
    const float det = bxx * (byy * bzz - byz * byz) - bxy * (bxy * bzz - byz * bxz) + bxz * (bxy * byz - byy * bxz);
    const float r = fminf(1.0f, fmaxf(-1.0f, 0.5f * det));
    return -p + 2.0f * q * cosf(acosf(r) * (1.0f / 3.0f));
}

// The union records a neighbour pass stages per round (S, one LDS buffer per wavefront) is a
// template parameter: every value gives the same results, and the launcher takes the one the
// context's tuning names.
template <bool STRENGTH, int S>
__global__ __launch_bounds__(kBlock) void kPass1(DeviceCtx c) {
    struct Rec {
        Vec4 p, v;
    };
    __shared__ Rec stageBuf[kWavesPerBlock][S];
    if (stepHeld(c)) return;
    const uint32_t b = xcdBlock(blockIdx.x, gridDim.x);
    const uint32_t i = b * kBlock + threadIdx.x;
    const uint32_t sg = c.seg.slotSeg[i];
    const bool active = i - c.seg.off[sg] < c.seg.n[sg];
    if (ballot(active) == 0ull) return;  // a wave of padding slots

    const float4 pi = c.cur.pos[i];
    const float4 vi = c.cur.vel[i];
    float4 hu = c.cur.hu[i];
    const uint32_t meta = c.cur.meta[i];
    const int mat = (int)(meta & 0xffu);
    const float hi = fmaxf(hu.x, 1.0e-6f), h0i = fmaxf(hu.y, 1.0e-6f), rhoI = fmaxf(vi.w, 1.0e-9f);
    const float Hi = 2.0f * hi;
    const float Hi2 = Hi * Hi;
    // f_surf takes 1 - C and |n| / h against the 0.15 threshold, which multiplies their errors by
    // up to 1.6 C / f_surf (hundreds, at a layer boundary inside a body where C dips to 0.9).  The
    // Shepard terms are therefore evaluated in FP64 from the FP32 inputs (the pair's separation
    // exactly, the kernel polynomial, the volume correctly rounded once per staged record), and
    // the Shepard sum and the normal accumulate in FP64.  wNorm applies once, at the end.
    const double pxd = (double)pi.x, pyd = (double)pi.y, pzd = (double)pi.z;
    const double invHd = 1.0 / (double)Hi;
    double shepW = 0.0;
    uint32_t ngb = 1;
    double nsx = 0.0, nsy = 0.0, nsz = 0.0;
    float gxx = 0.0f, gxy = 0.0f, gxz = 0.0f, gyx = 0.0f, gyy = 0.0f, gyz = 0.0f, gzx = 0.0f, gzy = 0.0f, gzz = 0.0f;

    const Support sI = supportOf(Hi);
    forEachNeighbour<S>(
        c.nb, i, stageBuf[threadIdx.x >> 6],
        [&](uint32_t j, Rec& r) {
            r.p = loadVec(c.cur.pos, j);
            r.v = loadVec(c.cur.vel, j);
        },
        [&](Rec& r) { r.v.w = __fdiv_rn(r.p.w, fmaxf(r.v.w, 1.0e-9f)); },  // V_j = m_j / rho_j, once per record
        [&](const Rec& n) {
            const float4 pj = toFloat4(n.p), vj = toFloat4(n.v);
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const float r2 = dist2f(dx, dy, dz);
            if (!(r2 < Hi2)) return;
            ++ngb;
            const float r = fastSqrt(r2);
            const float volJ = vj.w;
            const double ex = pxd - (double)pj.x, ey = pyd - (double)pj.y, ez = pzd - (double)pj.z;
            const double r2d = fma(ez, ez, fma(ey, ey, ex * ex));
            // r = r2 / sqrt(r2): the hardware FP32 reciprocal square root (about 1 ulp) refined by one
            // FP64 Newton step (relative error ~1e-14); a pair at zero separation has q = 0
            const double y0 = (double)__builtin_amdgcn_rsqf(r2);
            const double y1 = y0 * fma(-0.5 * r2d, y0 * y0, 1.5);
            const double qd = r2 > 0.0f ? r2d * y1 * invHd : 0.0;
            const double od = fmax(0.0, 1.0 - qd), o2d = od * od;
            const double vw = (double)volJ * (o2d * o2d) * fma(4.0, qd, 1.0);
            shepW += vw;
            nsx -= vw * ex;
            nsy -= vw * ey;
            nsz -= vw * ez;
            const float gc = volJ * wendlandGradS(r, sI);
            const float gx = gc * dx, gy = gc * dy, gz = gc * dz;
            const float dvx = vj.x - vi.x, dvy = vj.y - vi.y, dvz = vj.z - vi.z;
            gxx += dvx * gx; gxy += dvx * gy; gxz += dvx * gz;
            gyx += dvy * gx; gyy += dvy * gy; gyz += dvy * gz;
            gzx += dvz * gx; gzy += dvz * gy; gzz += dvz * gz;
        });
    if (!active) return;

    const double wNormD = 21.0 / (2.0 * 3.14159265358979323846) * invHd * invHd * invHd;
    const double shep = ((double)pi.w / (double)rhoI + shepW) * wNormD;
    const double normNs = wNormD * sqrt(nsx * nsx + nsy * nsy + nsz * nsz);
    float fs = (float)fmin(1.0, fmax(0.0, 1.6 * fmax(1.0 - shep, normNs / (double)hi) - 0.15));
    if (ngb <= 2u) fs = 1.0f;
    const float divV = gxx + gyy + gzz;
    const float cux = gzy - gyz, cuy = gxz - gzx, cuz = gyx - gxy;
    // the factor in FP64, rounded once: (float)cbrt(56/n) is the correctly rounded cube root for
    // every n where the clamp does not bind, so the oracle reproduces h' bit for bit
    const float factor = fminf(1.08f, fmaxf(0.92f, (float)cbrt(56.0 / (double)ngb)));
    const float hNew = fminf(6.0f * h0i, fmaxf(0.35f * h0i, hi * factor));
    hu.x = hNew;
    c.cur.hu[i] = hu;

    const EosResult e = evaluateEos(mat, (double)vi.w, (double)hu.z, (double)hu.w, (double)c.prm.gammaHHe);
    const float P = (float)e.p, cs = (float)e.cs, T = (float)e.T;
    c.d.eos[i] = make_float4(P, cs, T, 0.0f);
    c.d.phase[i] = e.phase;

    const float csF = fmaxf(cs, 0.05f);
    const float fBal = balsaraSwitch(divV, cux, cuy, cuz, csF, fmaxf(hNew, 1.0e-6f));
    const float invR2 = 1.0f / (rhoI * rhoI);
    const float pEff = effectivePressure(P, e.phase);
    c.d.nb[i] = make_float4(hNew, pEff * invR2, csF, fBal);
    c.d.p1a[i] = make_float4((float)shep, fs, divV, (float)ngb);
    c.d.p1b[i] = make_float4(cux, cuy, cuz, fBal);
    if (STRENGTH) {
        const float4 s0 = c.cur.str0[i];
        const float4 s1 = c.cur.str1[i];
        c.d.sr0[i] = make_float4(s0.x * invR2, s0.y * invR2, s0.z * invR2, s0.w * invR2);
        c.d.sr1[i] = make_float2(s1.x * invR2, s1.y * invR2);
    }
}

// EOS of the current state, and the pass-2 record of the particle refreshed with it: pass 2
// reads P_eff / rho^2, max(c_s, 0.05), the Balsara factor (from pass 1's div v and curl v and the
// new c_s) and S / rho^2 from the record, so an EOS evaluation followed by pass 2 without a pass 1
// in between (the relaxation's forces after its corrections) sees the new state.
template <bool STRENGTH>
__global__ __launch_bounds__(kBlock) void kEos(DeviceCtx c) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t sg = c.seg.slotSeg[i];
    if (i - c.seg.off[sg] >= c.seg.n[sg]) return;
    const float4 vi = c.cur.vel[i];
    const float4 hu = c.cur.hu[i];
    const int mat = (int)(c.cur.meta[i] & 0xffu);
    const EosResult e = evaluateEos(mat, (double)vi.w, (double)hu.z, (double)hu.w, (double)c.prm.gammaHHe);
    const float P = (float)e.p, cs = (float)e.cs;
    c.d.eos[i] = make_float4(P, cs, (float)e.T, 0.0f);
    c.d.phase[i] = e.phase;

    const float4 a = c.d.p1a[i];
    const float4 b = c.d.p1b[i];
    const float csF = fmaxf(cs, 0.05f);
    const float fBal = balsaraSwitch(a.z, b.x, b.y, b.z, csF, fmaxf(hu.x, 1.0e-6f));
    const float rhoI = fmaxf(vi.w, 1.0e-9f), invR2 = 1.0f / (rhoI * rhoI);
    c.d.nb[i] = make_float4(hu.x, effectivePressure(P, e.phase) * invR2, csF, fBal);
    c.d.p1b[i] = make_float4(b.x, b.y, b.z, fBal);
    if (STRENGTH) {
        const float4 s0 = c.cur.str0[i];
        const float4 s1 = c.cur.str1[i];
        c.d.sr0[i] = make_float4(s0.x * invR2, s0.y * invR2, s0.z * invR2, s0.w * invR2);
        c.d.sr1[i] = make_float2(s1.x * invR2, s1.y * invR2);
    }
}

// Pass-2 neighbour record staged in LDS: 64 bytes, 80 with the stress deviator S/rho^2 (a
// smaller record lets five blocks share a CU instead of four).  v.w holds max(rho_j, 1e-9) and
// volJ = m_j / rho_j (hardware reciprocal), once per staged record rather than per pair.
template <bool STRENGTH>
struct Pass2Rec {
    Vec4 p, n, v;  // n.x holds max(h_j, 1e-6)
    uint32_t mat;
    float invH, gNorm;  // the kernel gradient's constants of H_j, once per staged record
    float volJ;
};
template <>
struct Pass2Rec<true> {
    Vec4 p, n, v, s0;
    Vec2 s1;
    uint32_t mat;
    float volJ;
};
static_assert(sizeof(Pass2Rec<false>) == 64 && sizeof(Pass2Rec<true>) == 80, "pass-2 records are 4 or 5 quads");

template <bool STRENGTH, int S>
__global__ __launch_bounds__(kBlock) void kPass2(DeviceCtx c) {
    using Rec = Pass2Rec<STRENGTH>;
    __shared__ Rec stageBuf[kWavesPerBlock][S];
    if (stepHeld(c)) return;
    const uint32_t b = xcdBlock(blockIdx.x, gridDim.x);
    const uint32_t i = b * kBlock + threadIdx.x;
    const uint32_t sg = c.seg.slotSeg[i];
    const bool active = i - c.seg.off[sg] < c.seg.n[sg];
    if (ballot(active) == 0ull) return;  // a wave of padding slots

    const float4 pi = c.cur.pos[i];
    const float4 vi = c.cur.vel[i];
    const float4 nbi = c.d.nb[i];
    const float4 hu = c.cur.hu[i];
    const float4 eo = c.d.eos[i];
    const uint8_t phI = c.d.phase[i];
    const uint32_t meta = c.cur.meta[i];
    const uint32_t matI = meta & 0xffu;
    const float hi = fmaxf(nbi.x, 1.0e-6f), Hi = 2.0f * hi;
    const float rhoI = fmaxf(vi.w, 1.0e-9f), invR2 = 1.0f / (rhoI * rhoI);
    const float pI = effectivePressure(eo.x, phI);
    const float csI = nbi.z, fBalI = nbi.w;
    const float Ri = pI < 0.0f ? 0.3f * (-pI) * invR2 : 0.0f;
    const float pr2I = pI * invR2;
    float4 sI0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), sI1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 srI0 = sI0;
    float2 srI1 = make_float2(0.0f, 0.0f);
    if (STRENGTH) {
        sI0 = c.cur.str0[i];
        sI1 = c.cur.str1[i];
        srI0 = c.d.sr0[i];
        srI1 = c.d.sr1[i];
    }

    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    float dCont = 0.0f, dDiff = 0.0f, duP = 0.0f, duAv = 0.0f, duT = 0.0f, duS = 0.0f;
    float gxx = 0.0f, gxy = 0.0f, gxz = 0.0f, gyx = 0.0f, gyy = 0.0f, gyz = 0.0f, gzx = 0.0f, gzy = 0.0f, gzz = 0.0f;
    uint32_t pairs = 0;

    const Support sI = supportOf(Hi);
    const float dampI = 0.01f * Hi * Hi;
    forEachNeighbour<S>(
        c.nb, i, stageBuf[threadIdx.x >> 6],
        [&](uint32_t j, Rec& r) {
            r.p = loadVec(c.cur.pos, j);
            r.n = loadVec(c.d.nb, j);
            r.v = loadVec(c.cur.vel, j);
            r.mat = c.cur.meta[j];
            if constexpr (STRENGTH) {
                r.s0 = loadVec(c.d.sr0, j);
                r.s1 = loadVec(c.d.sr1, j);
            }
        },
        [&](Rec& r) {
            r.mat &= 0xffu;
            r.v.w = maxPos(r.v.w, 1.0e-9f);
            r.volJ = r.p.w * fastRcp(r.v.w);
            if constexpr (!STRENGTH) {
                r.n.x = maxPos(r.n.x, 1.0e-6f);
                const Support sj = supportOf(2.0f * r.n.x);
                r.invH = sj.invH;
                r.gNorm = sj.gNorm;
            }
        },
        [&](const Rec& n) {
            const float4 pj = toFloat4(n.p), nbj = toFloat4(n.n), vj = toFloat4(n.v);
            const float hj = STRENGTH ? maxPos(nbj.x, 1.0e-6f) : nbj.x, Hj = 2.0f * hj;
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const float r2 = dist2f(dx, dy, dz);
            const float Hm = fmaxf(Hi, Hj);
            if (!(r2 < Hm * Hm)) return;
            ++pairs;
            const float mj = pj.w, rhoJ = vj.w;
            const float pr2J = nbj.y, csJ = nbj.z, fBalJ = nbj.w;
            const float dvx = vi.x - vj.x, dvy = vi.y - vj.y, dvz = vi.z - vj.z;
            const float vDotR = dvx * dx + dvy * dy + dvz * dz;
            const float r = fastSqrt(r2);
            Support sJ;
            if constexpr (STRENGTH) sJ = supportOf(Hj);
            else sJ = Support{n.invH, 0.0f, n.gNorm};
            const float gI = wendlandGradS(r, sI), gJ = wendlandGradS(r, sJ);
            const float gb = 0.5f * (gI + gJ);
            const float gvx = gb * dx, gvy = gb * dy, gvz = gb * dz;
            const float vdg = dvx * gvx + dvy * gvy + dvz * gvz;
            const float volJ = n.volJ;
            dCont += volJ * vdg;
            if (n.mat == matI) {
                const float rdg = -(dx * gvx + dy * gvy + dz * gvz);
                dDiff += volJ * (2.0f * (rhoJ - rhoI) * fastRcp(r2 + dampI)) * rdg;
            }
            float piIJ = 0.0f;
            if (vDotR < 0.0f) {
                const float hBar = 0.5f * (hi + hj), cBar = 0.5f * (csI + csJ), rhoBar = 0.5f * (rhoI + rhoJ);
                const float mu = (hBar * vDotR) * fastRcp(r2 + 0.01f * hBar * hBar);
                piIJ = ((-cBar * mu + 2.0f * mu * mu) * fastRcp(rhoBar)) * (0.5f * (fBalI + fBalJ));
            }
            float tens = 0.0f;
            const float Rj = pr2J < 0.0f ? -0.3f * pr2J : 0.0f;
            if (Ri + Rj > 0.0f) {
                const float Hbar = 0.5f * (Hi + Hj);
                if (r < Hbar) {
                    const float q = r * fastRcp(Hbar), o = fmaxf(0.0f, 1.0f - q), o2 = o * o;
                    const float fr = (o2 * o2 * (1.0f + 4.0f * q)) * (1.0f / 0.6328125f);
                    const float fr2 = fr * fr;
                    tens = (Ri + Rj) * (fr2 * fr2);
                }
            }
            const float iso = -mj * (pr2I + pr2J + piIJ + tens);
            ax += iso * gvx;
            ay += iso * gvy;
            az += iso * gvz;
            duP += mj * vdg;
            duAv += 0.5f * mj * piIJ * vdg;
            duT += 0.5f * mj * tens * vdg;
            if constexpr (STRENGTH) {
                const float4 sj0 = toFloat4(n.s0);
                const float2 sj1 = toFloat2(n.s1);
                const float sxx = srI0.x + sj0.x, syy = srI0.y + sj0.y, szz = srI0.z + sj0.z;
                const float sxy = srI0.w + sj0.w, sxz = srI1.x + sj1.x, syz = srI1.y + sj1.y;
                ax += mj * (sxx * gvx + sxy * gvy + sxz * gvz);
                ay += mj * (sxy * gvx + syy * gvy + syz * gvz);
                az += mj * (sxz * gvx + syz * gvy + szz * gvz);
                const float wx = -dvx, wy = -dvy, wz = -dvz;
                duS += mj * ((sI0.x * wx + sI0.w * wy + sI1.x * wz) * gvx + (sI0.w * wx + sI0.y * wy + sI1.y * wz) * gvy +
                             (sI1.x * wx + sI1.y * wy + sI0.z * wz) * gvz);
                const float gx = gI * dx, gy = gI * dy, gz = gI * dz;
                gxx += volJ * wx * gx; gxy += volJ * wx * gy; gxz += volJ * wx * gz;
                gyx += volJ * wy * gx; gyy += volJ * wy * gy; gyz += volJ * wy * gz;
                gzx += volJ * wz * gx; gzy += volJ * wz * gy; gzz += volJ * wz * gz;
            }
        });
    if (!active) return;

    const float4 g = c.d.grav[i];
    const float4 p1 = c.d.p1a[i];
    const float fs = p1.y;
    const float diffMask = p1.x > 0.85f ? (1.0f - fs) : 0.0f;
    const float drho = rhoI * dCont + 0.05f * Hi * csI * diffMask * dDiff;
    const float T = eo.z, D = hu.w;

    // surface radiative cooling in FP64 (sph_gravity.cu:1219-1235)
    const bool living = c.ss.living[sg] != 0u;
    const uint32_t body = (meta >> 8) & 0xffu;
    const double tAmb = (living && body == 0u && D < 0.8f) ? 255.0 : 3.0;
    double pRad = 0.0;
    if (T > c.prm.coolMinT) {
        const double mSi = fmax((double)pi.w * 1.0e24, 1.0), rhoSi = fmax((double)rhoI * 1.0e6, 1.0e-3);
        const double rp = cbrt((3.0 * mSi) / (4.0 * 3.14159265358979323846 * rhoSi));
        const double area = 3.14159265358979323846 * rp * rp;
        const double Td = (double)T;
        const double dT4 = fmax(0.0, Td * Td * Td * Td - tAmb * tAmb * tAmb * tAmb);
        pRad = (double)fs * 0.9 * 5.670374419e-8 * area * dT4;
    }
    const float duRad = (float)(-pRad / (fmax((double)pi.w, 1.0e-30) * 1.0e27));
    const float du = pr2I * duP + duAv + duT + invR2 * duS + duRad;

    float dDc = 0.0f;
    if (STRENGTH) {
        const Strength st = strengthOf((int)matI, c.prm.strengthMask);
        float4 ds0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float2 ds1 = make_float2(0.0f, 0.0f);
        if (st.mu > 0.0f && st.tMelt > 0.0f) {
            const float tn = T / st.tMelt;
            const float muEff = (1.0f - D) * st.mu * fmaxf(0.0f, 1.0f - tn * tn);
            const float tr = (gxx + gyy + gzz) * (1.0f / 3.0f);
            const float exx = gxx - tr, eyy = gyy - tr, ezz = gzz - tr;
            const float exy = 0.5f * (gxy + gyx), exz = 0.5f * (gxz + gzx), eyz = 0.5f * (gyz + gzy);
            const float rxy = 0.5f * (gxy - gyx), rxz = 0.5f * (gxz - gzx), ryz = 0.5f * (gyz - gzy);
            const float ryx = -rxy, rzx = -rxz, rzy = -ryz;
            const float sxx = sI0.x, syy = sI0.y, szz = sI0.z, sxy = sI0.w, sxz = sI1.x, syz = sI1.y;
            ds0.x = 2.0f * muEff * exx + 2.0f * (sxy * rxy + sxz * rxz);
            ds0.y = 2.0f * muEff * eyy + 2.0f * (sxy * ryx + syz * ryz);
            ds0.z = 2.0f * muEff * ezz + 2.0f * (sxz * rzx + syz * rzy);
            ds0.w = 2.0f * muEff * exy + sxx * ryx + sxz * ryz + syy * rxy + syz * rxz;
            ds1.x = 2.0f * muEff * exz + sxx * rzx + sxy * rzy + szz * rxz + syz * rxy;
            ds1.y = 2.0f * muEff * eyz + sxy * rzx + syy * rzy + sxz * ryx + szz * ryz;
            if (T < st.tMelt && D < 1.0f) {
                const float sm = maxPrincipalStress(pI, sxx, syy, szz, sxy, sxz, syz);
                const float stiff = fmaxf(st.bulk + (4.0f / 3.0f) * st.mu, 1.0e-12f);
                const float eT = fmaxf(0.0f, sm) / stiff;
                const float eF = fmaxf(sI1.w, 1.0e-8f);
                const float cCrack = fmaxf(csI, sqrtf((st.bulk + (4.0f / 3.0f) * st.mu) / rhoI));
                dDc = eT > eF ? 0.4f * cCrack / hi : 0.0f;
            }
        }
        c.d.dS0[i] = ds0;
        c.d.dS1[i] = ds1;
    }
    c.d.rate[i] = make_float4(g.x + ax, g.y + ay, g.z + az, drho);
    c.d.rate2[i] = make_float4(du, dDc, (float)pRad, (float)pairs);
}

template <int S>
void pass1As(const DeviceCtx& c, hipStream_t s) {
    if (c.strength) kPass1<true, S><<<c.blocks, kBlock, 0, s>>>(c);
    else kPass1<false, S><<<c.blocks, kBlock, 0, s>>>(c);
}
template <int S>
void pass2As(const DeviceCtx& c, hipStream_t s) {
    if (c.strength) kPass2<true, S><<<c.blocks, kBlock, 0, s>>>(c);
    else kPass2<false, S><<<c.blocks, kBlock, 0, s>>>(c);
}

}  // namespace

void launchPass1(const DeviceCtx& c, hipStream_t s) {
    if (c.tune.pass1Round == 128u) pass1As<128>(c, s);
    else pass1As<256>(c, s);
}

void launchEos(const DeviceCtx& c, hipStream_t s) {
    if (c.strength) kEos<true><<<c.blocks, kBlock, 0, s>>>(c);
    else kEos<false><<<c.blocks, kBlock, 0, s>>>(c);
}

void launchPass2(const DeviceCtx& c, hipStream_t s) {
    if (c.tune.pass2Round == 160u) pass2As<160>(c, s);
    else pass2As<128>(c, s);
}

}  // namespace astraea_hip
