// oracle.cpp — FP64 CPU restatement of the reference's per-step physics path.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the timed CPU
// baseline.  Never linked or loaded by the product.
//
// Every function cites the reference lines it restates (paths relative to the
// reference repository).  Build: oracle/Makefile (g++ -O2 -fopenmp
// -ffp-contract=off, so the FP32 decision replicas round exactly as written).
//
// Parity status (how each function is pinned; see tests/):
//   EOS (eos_*)                  closed forms of tests/cpp/test_eos.cpp:26-56 (u_cold, P_c, P_e,
//                                two-phase blend), H/He identity, phase branches, STP zero pressure.
//   morton_key / sort / karras   hand-checked golden vectors (tests/golden/), brute-force stable
//                                sort, tree invariants (every node one parent, leaf ranges partition).
//   moments                      root mass = sum m, root COM = system COM, boxes contain children.
//   gravity                      brute-force FP64 N^2 (reference AC3 thresholds,
//                                tests/cpp/test_gravity_sph.cpp:293-395) and theta -> 0 limit = N^2.
//   pass1                        linear velocity fields on a lattice: grad v = A M_i from brute-force
//                                lattice sums, M_i = identity to lattice accuracy; Shepard sums and
//                                f_surf by brute force (interior 0, free surface > 0); the h update on
//                                configurations of known neighbour count, with both clamps.
//   pass2                        isolated pairs: Monaghan viscosity x Balsara (with and without curl),
//                                tensile term = R_i + R_j at a quarter of the mean support, stress
//                                force and work (pair energy conserved); lattice: delta-SPH as a
//                                Laplacian estimate (2 kappa rho0) restricted to one material, Jaumann
//                                rate = 2 mu' strain + W S - S W (rigid rotation: co-rotation only);
//                                Grady-Kipp rate at the failure strain (numpy principal stresses);
//                                cooling: isolated closed form (tests/cpp/test_gravity_sph.cpp:646-670),
//                                255 K ambient of an intact living body, the threshold knob.
//   drift                        XSPH: zero for linear fields, 0.25 kappa H^2/15 for v = kappa y^2.
//   kick / dt                    closed form per particle; von Mises radial return by hand.
//   (tests/test_oracle_mutants.py builds copies with a constant, sign or index of these changed and
//    checks that the pinning tests fail on each.)
//   step                         conservation gates of engine/headless_main.cpp:263-281.
//   relax                        closed forms: steps = 0 gives v_rms / v_esc of the loaded state;
//                                force-free particles keep 0.85^steps of their velocity about the
//                                centre of mass through the undamped steps; u after the last
//                                re-anchor equals u_cold(rho) + c_v T_anchor; stress and damage zero;
//                                centre of mass and momentum of a real body stay at zero.
//   classify_arrays              hand-solvable configurations: a test particle on circular,
//                                parabolic-plus and eccentric orbits about a point remnant (disk /
//                                ejecta / planet by the vis-viva and periapsis closed forms), rigid
//                                rotation (spin = omega), the seed at the phi minimum, first-index
//                                ties, and the stop rule when nothing is bound.

#include "oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kG = 66.7430;             // G_SI * M * T^2 / L^3, engine/units.hpp:26
constexpr float kGf = 66.7430f;
constexpr double kMUnit = 1.0e24, kRhoUnit = 1.0e6, kPUnit = 1.0e12, kUUnit = 1.0e6;
constexpr double kSigmaSB = 5.670374419e-8;  // engine/units.hpp:8

// ---------------------------------------------------------------------------
// Material table: engine/eos.cu:10-116 (Tillotson / H-He parameters) and the
// strength table engine/sph_gravity.cu:42-73 (same numbers, SI units).
// ---------------------------------------------------------------------------
struct Mat {
    double rho0, a, b, A, B, E0, Eiv, Ecv, alpha, beta, cv, tMelt, tRef, mu, Y0, epsMin, kH2, gamma;
    double bulk;  // bulk modulus of the strength table, sph_gravity.cu:42-73
};
const Mat kMat[5] = {
    {7800.0, 0.5, 1.5, 1.28e11, 1.05e11, 9.50e6, 2.40e6, 8.67e6, 5.0, 5.0, 449.0, 1811.0, 293.0, 7.70e10, 6.80e8, 1.0e-3, 0.0, 1.0, 1.28e11},
    {2680.0, 0.5, 1.3, 1.80e10, 1.80e10, 1.60e7, 3.50e6, 1.80e7, 5.0, 5.0, 790.0, 1600.0, 293.0, 2.20e10, 3.50e9, 1.0e-4, 0.0, 1.0, 1.80e10},
    {1000.0, 0.7, 0.15, 2.18e9, 1.325e10, 7.00e6, 4.19e5, 2.68e6, 10.0, 5.0, 4184.0, 273.15, 273.15, 3.50e9, 1.00e7, 5.0e-4, 0.0, 1.0, 2.18e9},
    {0.178, 0.4, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 8260.0, 0.0, 0.0, 0.0, 0.0, 1.0, 2.08e5, 1.4, 0.0},
    {2000.0, 0.5, 1.3, 1.20e10, 1.20e10, 1.60e7, 3.50e6, 1.80e7, 5.0, 5.0, 800.0, 1500.0, 293.0, 8.00e9, 5.00e7, 5.0e-5, 0.0, 1.0, 1.20e10},
};
inline const Mat& matOf(int m) { return kMat[(m >= 0 && m <= 4) ? m : 1]; }
inline int matIdx(int m) { return (m >= 0 && m <= 4) ? m : 1; }

enum { kSolid = 0, kMolten = 1, kVapor = 2, kPlasma = 3, kLiquidWater = 4, kGas = 5 };

// u_cold(rho): engine/eos.cu:283-318
double coldEnergy(int mat, double rho) {
    const Mat& p = matOf(mat);
    rho = std::max(rho, 0.0);
    if (matIdx(mat) == 3) return p.kH2 * rho;
    const double rho0 = p.rho0 / kRhoUnit;
    const double eta = rho / rho0;
    if (eta <= 1.0) return 0.0;
    const double mu = eta - 1.0, A = p.A / kPUnit, B = p.B / kPUnit;
    double tA, tB;
    if (mu < 1.0e-5) {
        const double m2 = mu * mu, m3 = m2 * mu, m4 = m3 * mu, m5 = m4 * mu;
        tA = 0.5 * m2 - (2.0 / 3.0) * m3 + 0.75 * m4 - 0.8 * m5;
        tB = (1.0 / 3.0) * m3 - 0.5 * m4 + 0.6 * m5;
    } else {
        const double lg = std::log1p(mu);
        tA = lg - mu / eta;
        tB = mu * (2.0 + mu) / eta - 2.0 * lg;
    }
    return (A * tA + B * tB) / rho0;
}

struct PDeriv { double p, dpdrho, dpdu; bool damagedTension; };

// Compressed Tillotson branch: engine/eos.cu:136-180
PDeriv tillotsonCompressed(const Mat& p, double rho, double rho0, double eta, double mu, double u, double dmg) {
    const double A = p.A / kPUnit, B = p.B / kPUnit, E0 = p.E0 / kUUnit;
    const double bEff = mu >= 0.0 ? B : 0.0;
    const double den = E0 * eta * eta;
    const double ratio = den > 1.0e-30 ? u / den : 1.0e30;
    const double invW = 1.0 / (ratio + 1.0);
    const double wm1w2 = ratio > 1.0e15 ? 0.0 : ratio * invW * invW;
    PDeriv r;
    r.p = (p.a + p.b * invW) * rho * u + A * mu + bEff * mu * mu;
    r.dpdrho = (p.a + p.b * invW) * u + 2.0 * p.b * u * wm1w2 + (A + 2.0 * bEff * mu) / rho0;
    r.dpdu = (p.a + p.b * invW * invW) * rho;
    r.damagedTension = false;
    if (r.p < 0.0 && dmg > 0.0) {
        const double k = 1.0 - dmg;
        r.p *= k; r.dpdrho *= k; r.dpdu *= k;
        r.damagedTension = true;
    }
    return r;
}

// Expanded Tillotson branch: engine/eos.cu:182-227
PDeriv tillotsonExpanded(const Mat& p, double rho, double rho0, double eta, double mu, double u) {
    const double A = p.A / kPUnit, E0 = p.E0 / kUUnit;
    const double x = 1.0 / eta - 1.0;
    const double ax2 = p.alpha * x * x, bx = p.beta * x;
    const double eA = ax2 < 700.0 ? std::exp(-ax2) : 0.0;
    PDeriv r{};
    if (eA == 0.0) {
        r.p = p.a * rho * u; r.dpdrho = p.a * u; r.dpdu = p.a * rho;
        return r;
    }
    const double eB = bx < 700.0 ? std::exp(-bx) : 0.0;
    const double den = E0 * eta * eta;
    const double ratio = den > 1.0e-30 ? u / den : 1.0e30;
    const double invW = 1.0 / (ratio + 1.0);
    const double wm1w2 = ratio > 1.0e15 ? 0.0 : ratio * invW * invW;
    const double ideal = p.a * rho * u;
    const double inner = p.b * rho * u * invW + A * mu * eB;
    r.p = ideal + inner * eA;
    const double r0r2 = rho0 / (rho * rho);
    const double expDrho = eB > 0.0 ? (A / rho0 + A * mu * p.beta * r0r2) * eB : 0.0;
    r.dpdrho = p.a * u + eA * (p.b * u * invW + 2.0 * p.b * u * wm1w2 + expDrho) + 2.0 * p.alpha * x * r0r2 * (r.p - ideal);
    r.dpdu = (p.a + p.b * eA * invW * invW) * rho;
    return r;
}

// Phase classification: engine/eos.cu:229-267
int classifyPhase(int mat, const Mat& p, double rho, double u, double T) {
    const int mi = matIdx(mat);
    if (mi == 3) return T > 10000.0 ? kPlasma : kGas;
    const double rhoSI = rho * kRhoUnit, ecv = p.Ecv / kUUnit;
    if (mi == 2) {
        if (T > 15000.0) return kPlasma;
        if (u >= ecv || (T > 647.0 && rhoSI < 322.0)) return kVapor;
        if (T >= 273.15) return kLiquidWater;
        return kSolid;
    }
    if (T > 12000.0) return kPlasma;
    if (u >= ecv && rhoSI < 0.8 * p.rho0) return kVapor;
    if (T >= p.tMelt) return kMolten;
    return kSolid;
}

struct EosOut { double p, cs, T; int phase; };

// evaluateEosPsu: engine/eos.cu:320-388; gammaHHe replaces the table's 1.4 (eos.cu:93)
EosOut eosEval(int mat, double rhoIn, double u, double dmgIn, double gammaHHe) {
    const Mat& p = matOf(mat);
    const double rho = std::max(rhoIn, 1.0e-12);
    const double dmg = std::min(std::max(dmgIn, 0.0), 1.0);
    const double cv = p.cv / kUUnit;
    const double uc = coldEnergy(mat, rho);
    EosOut o{};
    if (matIdx(mat) == 3) {
        const double uTh = std::max(0.0, u - uc);
        const double T = std::max(3.0, uTh / cv);
        o.p = p.kH2 * rho * rho + (gammaHHe - 1.0) * rho * uTh;
        const double cs2 = 2.0 * p.kH2 * rho + gammaHHe * (gammaHHe - 1.0) * uTh;
        o.cs = std::max(0.05, std::sqrt(std::max(0.0, cs2)));
        o.T = T;
        o.phase = classifyPhase(mat, p, rho, u, T);
        return o;
    }
    const double rho0 = p.rho0 / kRhoUnit, eta = rho / rho0, mu = eta - 1.0;
    const double T = std::max(3.0, (u - uc) / cv);
    const double uE = uc + std::max(0.0, u - uc - cv * p.tRef);
    const double Eiv = p.Eiv / kUUnit, Ecv = p.Ecv / kUUnit, A = p.A / kPUnit;
    PDeriv r;
    bool dt = false;
    if (eta >= 1.0 || uE < Eiv) {
        r = tillotsonCompressed(p, rho, rho0, eta, mu, uE, dmg);
        dt = r.damagedTension;
    } else if (uE >= Ecv) {
        r = tillotsonExpanded(p, rho, rho0, eta, mu, uE);
    } else {
        const PDeriv c = tillotsonCompressed(p, rho, rho0, eta, mu, uE, dmg);
        const PDeriv e = tillotsonExpanded(p, rho, rho0, eta, mu, uE);
        dt = c.damagedTension;
        const double den = std::max(1.0e-30, Ecv - Eiv);
        const double f = std::min(std::max((uE - Eiv) / den, 0.0), 1.0);
        r.p = (1.0 - f) * c.p + f * e.p;
        r.dpdrho = (1.0 - f) * c.dpdrho + f * e.dpdrho;
        r.dpdu = (1.0 - f) * c.dpdu + f * e.dpdu + (e.p - c.p) / den;
    }
    const double cs2 = r.dpdrho + (r.p / (rho * rho)) * r.dpdu;
    const bool condensed = eta >= 1.0 || (uE < Eiv && rho >= 0.9 * rho0 && !dt);
    const double cMin = condensed ? std::max(0.05, std::sqrt(A / rho0)) : 0.05;
    o.p = r.p;
    o.cs = std::max(cMin, std::sqrt(std::max(0.0, cs2)));
    o.T = T;
    o.phase = classifyPhase(mat, p, rho, u, T);
    return o;
}

// evaluateFromRhoTempSI: engine/eos.cu:391-440
void eosFromRhoTemp(int mat, double rhoKg, double TK, double& uSi, double& pSi) {
    const Mat& p = matOf(mat);
    const double rho = std::max(rhoKg, 1.0e-12), Tc = std::max(TK, 0.0);
    if (matIdx(mat) == 3) {
        const double uc = p.kH2 * rho, uth = p.cv * Tc;
        uSi = uc + uth;
        pSi = p.kH2 * rho * rho + (p.gamma - 1.0) * rho * uth;
        return;
    }
    const double rhoP = rho / kRhoUnit, ucP = coldEnergy(mat, rhoP), ucSi = ucP * kUUnit;
    uSi = ucSi + p.cv * Tc;
    const double uE = ucP + std::max(0.0, uSi - ucSi - p.cv * p.tRef) / kUUnit;
    const double rho0 = p.rho0 / kRhoUnit, eta = rhoP / rho0, mu = eta - 1.0;
    const double Eiv = p.Eiv / kUUnit, Ecv = p.Ecv / kUUnit;
    double pp;
    if (eta >= 1.0 || uE < Eiv) pp = tillotsonCompressed(p, rhoP, rho0, eta, mu, uE, 0.0).p;
    else if (uE >= Ecv) pp = tillotsonExpanded(p, rhoP, rho0, eta, mu, uE).p;
    else {
        const double pc = tillotsonCompressed(p, rhoP, rho0, eta, mu, uE, 0.0).p;
        const double pe = tillotsonExpanded(p, rhoP, rho0, eta, mu, uE).p;
        const double f = std::min(std::max((uE - Eiv) / std::max(1.0e-30, Ecv - Eiv), 0.0), 1.0);
        pp = (1.0 - f) * pc + f * pe;
    }
    pSi = pp * kPUnit;
}

// invertDensityFromPressureTempSI: engine/eos.cu:442-534 (bracket + secant/bisection)
double invertDensity(int mat, double target, double TK, double guess) {
    const Mat& p = matOf(mat);
    if (matIdx(mat) == 3) {
        if (target <= 0.0) return 1.0e-6;
        const double bl = (p.gamma - 1.0) * p.cv * std::max(TK, 0.0);
        const double d = std::sqrt(bl * bl + 4.0 * p.kH2 * target);
        return 2.0 * target / std::max(1.0e-30, bl + d);
    }
    auto P = [&](double r) { double u, pp; eosFromRhoTemp(mat, r, TK, u, pp); return pp; };
    double r0 = guess > 0.0 ? guess : p.rho0;
    r0 = std::min(std::max(r0, 0.05 * p.rho0), 50.0 * p.rho0);
    double lo = r0, hi = r0;
    const double p0 = P(r0);
    if (p0 > target) {
        for (int i = 0; i < 64; ++i) {
            lo *= i < 12 ? 0.85 : 0.5;
            if (lo < 1.0e-6) lo = 1.0e-6;
            if (P(lo) <= target || lo <= 1.0e-6) break;
        }
    } else if (p0 < target) {
        for (int i = 0; i < 64; ++i) {
            hi *= i < 12 ? 1.25 : 2.0;
            if (P(hi) >= target || hi > 1.0e8) break;
        }
    } else {
        return r0;
    }
    double fl = P(lo) - target, fh = P(hi) - target;
    if (fl >= 0.0) return lo;
    if (fh <= 0.0) return hi;
    double r = 0.5 * (lo + hi);
    for (int it = 0; it < 64; ++it) {
        const double fm = P(r) - target;
        if (std::abs(fm) <= 1.0e-9 * std::max(1.0, std::abs(target))) return r;
        if (fm < 0.0) { lo = r; fl = fm; } else { hi = r; fh = fm; }
        if (hi - lo <= 1.0e-12 * std::max(1.0, r)) return 0.5 * (lo + hi);
        double nx = 0.5 * (lo + hi);
        const double den = fh - fl;
        if (std::abs(den) > 1.0e-30) {
            const double s = lo - fl * (hi - lo) / den;
            if (s > lo && s < hi) nx = s;
        }
        r = nx;
    }
    return 0.5 * (lo + hi);
}

// ---------------------------------------------------------------------------
// FP32 helpers replicated bit-for-bit by the HIP kernels.
// ---------------------------------------------------------------------------

// expandBits21 / morton3D64: engine/sph_gravity.cu:151-171
inline uint64_t spread21(uint64_t v) {
    v &= 0x1fffffULL;
    v = (v | (v << 32)) & 0x001f00000000ffffULL;
    v = (v | (v << 16)) & 0x001f0000ff0000ffULL;
    v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}
inline uint32_t quantise(float x, float mn, float inv) {
    const float f = std::fmin(std::fmax((x - mn) * inv, 0.0f), 2097151.0f);
    return static_cast<uint32_t>(f);
}
inline float invSpanOf(float mn, float mx) {
    const float span = mx - mn;
    return span > 1.0e-6f ? 2097151.0f / span : 0.0f;  // sph_gravity.cu:347-352
}
inline uint64_t mortonKey(float x, float y, float z, const float* b, const float* inv) {
    return (spread21(quantise(x, b[0], inv[0])) << 2) | (spread21(quantise(y, b[1], inv[1])) << 1) |
           spread21(quantise(z, b[2], inv[2]));
}
// The mixed form's Morton box (FP64 accumulators): the same quantisation in FP64, from the FP64
// positions and an FP64 box (the kernels' quantise64 / invSpan64).
inline uint32_t quantise64(double x, double mn, double inv) {
    const double f = std::fmin(std::fmax((x - mn) * inv, 0.0), 2097151.0);
    return static_cast<uint32_t>(f);
}
inline double invSpan64Of(double mn, double mx) {
    const double span = mx - mn;
    return span > 1.0e-6 ? 2097151.0 / span : 0.0;
}
inline uint64_t mortonKey64(double x, double y, double z, const double* b, const double* inv) {
    return (spread21(quantise64(x, b[0], inv[0])) << 2) | (spread21(quantise64(y, b[1], inv[1])) << 1) |
           spread21(quantise64(z, b[2], inv[2]));
}

// Squared distance with the kernels' operation order: ((dx*dx) + dy*dy) + dz*dz as two FMAs.
inline float dist2f(float dx, float dy, float dz) {
    float t = dx * dx;
    t = std::fmaf(dy, dy, t);
    t = std::fmaf(dz, dz, t);
    return t;
}

// karrasDelta: engine/sph_gravity.cu:173-187
inline int karrasDelta(const uint64_t* k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = k[i], b = k[j];
    if (a != b) return __builtin_clzll(a ^ b);
    const uint32_t x = static_cast<uint32_t>(i) ^ static_cast<uint32_t>(j);
    return 64 + (x == 0u ? 32 : __builtin_clz(x));  // CUDA/HIP __clz(0) == 32
}

// Support bound of one particle used by the neighbour-table build: the largest
// smoothing support 2*max(h',1e-6) the particle can have after pass 1's update
// h' = clamp(h*f, 0.35 h0, 6 h0), f in [0.92, 1.08] (sph_gravity.cu:785-787).
inline float supportBoundF(float h, float h0) {
    return 2.0f * std::fmax(std::fmax(std::fmax(h, 1.0e-6f) * 1.08f, 0.35f * std::fmax(h0, 1.0e-6f)), 1.0e-6f);
}
inline double supportBoundD(double h, double h0) {
    return 2.0 * std::max(std::max(std::max(h, 1.0e-6) * 1.08, 0.35 * std::max(h0, 1.0e-6)), 1.0e-6);
}

// Wendland C2 in 3D with support H: engine/sph_gravity.cu:199-220
inline double wendland(double r, double H) {
    if (r >= H || H <= 0.0) return 0.0;
    const double q = r / H, o = 1.0 - q, o2 = o * o;
    return 21.0 / (2.0 * kPi * H * H * H) * o2 * o2 * (1.0 + 4.0 * q);
}
inline double wendlandGrad(double r, double H) {
    if (r >= H || H <= 0.0) return 0.0;
    const double q = r / H, o = 1.0 - q, H2 = H * H;
    return -(210.0 / (kPi * H2 * H2 * H)) * o * o * o;
}
// computeBalsaraSwitch: engine/sph_gravity.cu:222-228
inline double balsara(double divv, double cx, double cy, double cz, double cs, double h) {
    const double ad = std::abs(divv), ac = std::sqrt(cx * cx + cy * cy + cz * cz);
    const double den = ad + ac + 1.0e-4 * std::max(cs, 0.05) / std::max(h, 1.0e-6);
    return ad / std::max(den, 1.0e-30);
}
// maxPrincipalStressFromTraceless: engine/sph_gravity.cu:230-254
inline double maxPrincipal(double P, const double* s) {
    const double p2 = s[0] * s[0] + s[1] * s[1] + s[2] * s[2] + 2.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
    if (p2 <= 1.0e-24) return -P;
    const double q = std::sqrt(p2 / 6.0), iq = 1.0 / q;
    const double bxx = s[0] * iq, byy = s[1] * iq, bzz = s[2] * iq, bxy = s[3] * iq, bxz = s[4] * iq, byz = s[5] * iq;
    const double det = bxx * (byy * bzz - byz * byz) - bxy * (bxy * bzz - byz * bxz) + bxz * (bxy * byz - byy * bxz);
    const double r = std::min(1.0, std::max(-1.0, 0.5 * det));
    return -P + 2.0 * q * std::cos(std::acos(r) / 3.0);
}

// ---------------------------------------------------------------------------
// State
// ---------------------------------------------------------------------------
struct Seg {
    uint32_t n = 0;
    // persistent
    std::vector<double> x, y, z, vx, vy, vz, m, h, h0, rho, u, s[6], D, Dc, eps;
    std::vector<uint8_t> mat, body, tracer;
    std::vector<uint32_t> oid;
    // derived
    std::vector<double> ax, ay, az, gx, gy, gz, phi, drho, du, P, cs, T, shep, fs, divv, cu[3], dS[6], dD, prad;
    std::vector<uint8_t> phase;
    std::vector<uint32_t> ngb;
    // tree of the last sort (local indexing as the reference)
    bool tree = false;
    std::vector<uint64_t> keys;
    std::vector<uint32_t> perm;
    std::vector<int32_t> left, right, parI, parL;
    std::vector<float> cm4, mn4, mx4;          // FP32 replica (com.m / min.extent / max.supportBound)
    std::vector<double> cmx, cmy, cmz, cmm;    // FP64 monopoles
    std::vector<double> bmn, bmx, bH;          // FP64 boxes (3 per node) and max support bound
    bool living = false;
    // integrator
    double dt = 0.0, prevDt = 0.0, erad = 0.0, prevRadW = 0.0, simTime = 0.0;
    uint64_t interactions = 0, overflows = 0, pairs1 = 0, pairs2 = 0;

    void resize(uint32_t nn) {
        n = nn;
        for (auto* v : {&x, &y, &z, &vx, &vy, &vz, &m, &h, &h0, &rho, &u, &s[0], &s[1], &s[2], &s[3], &s[4], &s[5], &D, &Dc,
                        &ax, &ay, &az, &gx, &gy, &gz, &phi, &drho, &du, &P, &cs, &T, &shep, &fs, &divv, &cu[0], &cu[1],
                        &cu[2], &dS[0], &dS[1], &dS[2], &dS[3], &dS[4], &dS[5], &dD, &prad})
            v->assign(nn, 0.0);
        eps.assign(nn, 1.0);
        mat.assign(nn, 0); body.assign(nn, 0); tracer.assign(nn, 0); phase.assign(nn, 0);
        oid.resize(nn);
        ngb.assign(nn, 0);
        tree = false;
    }
};

}  // namespace

struct or_ctx {
    astraea_params prm{};
    int G = 64;
    std::vector<Seg> seg;
};

namespace {

inline bool strengthOn(const or_ctx* c, int mat) {
    const int mi = matIdx(mat);
    const Mat& p = kMat[mi];
    return ((c->prm.strength_mask >> mi) & 1u) && p.mu > 0.0 && p.tMelt > 0.0;
}

// Apply a function to every field vector of a segment (used by the permutation).
template <class F>
void forEachPersistentAndDerived(Seg& s, F&& f) {
    for (auto* v : {&s.x, &s.y, &s.z, &s.vx, &s.vy, &s.vz, &s.m, &s.h, &s.h0, &s.rho, &s.u, &s.s[0], &s.s[1], &s.s[2],
                    &s.s[3], &s.s[4], &s.s[5], &s.D, &s.Dc, &s.eps, &s.ax, &s.ay, &s.az, &s.gx, &s.gy, &s.gz, &s.phi,
                    &s.drho, &s.du, &s.P, &s.cs, &s.T, &s.shep, &s.fs, &s.divv, &s.cu[0], &s.cu[1], &s.cu[2], &s.dS[0],
                    &s.dS[1], &s.dS[2], &s.dS[3], &s.dS[4], &s.dS[5], &s.dD, &s.prad})
        f(*v);
}

template <class T>
void permuteVec(std::vector<T>& v, const std::vector<uint32_t>& perm) {
    std::vector<T> o(v.size());
    for (size_t i = 0; i < perm.size(); ++i) o[i] = v[perm[i]];
    v.swap(o);
}

// launchSortAndBuildBvh: engine/sph_gravity.cu:1996-2061 (bounds 258-333, Morton 335-361,
// CUB stable sort 2028-2046, permutation 363-430, Karras 432-529, moments 531-620).
void sortAndBuild(or_ctx* c, Seg& s) {
    const uint32_t n = s.n;
    if (n == 0) return;
    std::vector<uint64_t> key(n);
    if (c->prm.precision == ASTRAEA_PREC_FP64_ACCUM) {
        double b[6] = {1.0e30, 1.0e30, 1.0e30, -1.0e30, -1.0e30, -1.0e30};
        for (uint32_t i = 0; i < n; ++i) {
            const double p[3] = {s.x[i], s.y[i], s.z[i]};
            for (int a = 0; a < 3; ++a) {
                b[a] = std::fmin(b[a], p[a]);
                b[3 + a] = std::fmax(b[3 + a], p[a]);
            }
        }
        const double inv[3] = {invSpan64Of(b[0], b[3]), invSpan64Of(b[1], b[4]), invSpan64Of(b[2], b[5])};
        for (uint32_t i = 0; i < n; ++i) key[i] = mortonKey64(s.x[i], s.y[i], s.z[i], b, inv);
    } else {
        float b[6] = {1.0e30f, 1.0e30f, 1.0e30f, -1.0e30f, -1.0e30f, -1.0e30f};
        for (uint32_t i = 0; i < n; ++i) {
            const float p[3] = {(float)s.x[i], (float)s.y[i], (float)s.z[i]};
            for (int a = 0; a < 3; ++a) {
                b[a] = std::fmin(b[a], p[a]);
                b[3 + a] = std::fmax(b[3 + a], p[a]);
            }
        }
        const float inv[3] = {invSpanOf(b[0], b[3]), invSpanOf(b[1], b[4]), invSpanOf(b[2], b[5])};
        for (uint32_t i = 0; i < n; ++i) key[i] = mortonKey((float)s.x[i], (float)s.y[i], (float)s.z[i], b, inv);
    }
    s.perm.resize(n);
    std::iota(s.perm.begin(), s.perm.end(), 0u);
    // Stable LSD radix sort over bits [0,63) == stable sort on the key (bit 63 is always 0).
    std::stable_sort(s.perm.begin(), s.perm.end(), [&](uint32_t a, uint32_t bb) { return key[a] < key[bb]; });
    s.keys.resize(n);
    for (uint32_t i = 0; i < n; ++i) s.keys[i] = key[s.perm[i]];
    forEachPersistentAndDerived(s, [&](std::vector<double>& v) { permuteVec(v, s.perm); });
    permuteVec(s.mat, s.perm); permuteVec(s.body, s.perm); permuteVec(s.tracer, s.perm);
    permuteVec(s.phase, s.perm); permuteVec(s.oid, s.perm); permuteVec(s.ngb, s.perm);

    // living flags: any tracer anchor with D < 0.8 (sph_gravity.cu:478-483)
    s.living = false;
    for (uint32_t i = 0; i < n; ++i)
        if (s.tracer[i] && (float)s.D[i] < 0.8f) s.living = true;

    const int ni = (int)n;
    const int nInt = n >= 2 ? ni - 1 : 0;
    s.left.assign(nInt, 0); s.right.assign(nInt, 0); s.parI.assign(nInt, -1); s.parL.assign(n, -1);
    const uint64_t* k = s.keys.data();
    for (int i = 0; i < nInt; ++i) {
        const int d = (karrasDelta(k, ni, i, i + 1) - karrasDelta(k, ni, i, i - 1)) >= 0 ? 1 : -1;
        const int dmin = karrasDelta(k, ni, i, i - d);
        int lmax = 2;
        while (karrasDelta(k, ni, i, i + lmax * d) > dmin) lmax <<= 1;
        int l = 0;
        for (int t = lmax >> 1; t >= 1; t >>= 1)
            if (karrasDelta(k, ni, i, i + (l + t) * d) > dmin) l += t;
        const int j = i + l * d;
        const int dnode = karrasDelta(k, ni, i, j);
        int sp = 0, t = l;
        do {
            t = (t + 1) >> 1;
            if (karrasDelta(k, ni, i, i + (sp + t) * d) > dnode) sp += t;
        } while (t > 1);
        const int gamma = i + sp * d + (d < 0 ? d : 0);
        const int first = std::min(i, j), last = std::max(i, j);
        s.left[i] = first == gamma ? ~gamma : gamma;
        s.right[i] = last == gamma + 1 ? ~(gamma + 1) : gamma + 1;
        if (first == gamma) s.parL[gamma] = i; else s.parI[gamma] = i;
        if (last == gamma + 1) s.parL[gamma + 1] = i; else s.parI[gamma + 1] = i;
    }

    // Moments: post-order over internal nodes (children complete before parents), the same
    // pairwise combine as the atomic-flag climb (sph_gravity.cu:531-620).
    s.cm4.assign(4 * (size_t)nInt, 0.0f); s.mn4.assign(4 * (size_t)nInt, 0.0f); s.mx4.assign(4 * (size_t)nInt, 0.0f);
    s.cmx.assign(nInt, 0.0); s.cmy.assign(nInt, 0.0); s.cmz.assign(nInt, 0.0); s.cmm.assign(nInt, 0.0);
    s.bmn.assign(3 * (size_t)nInt, 0.0); s.bmx.assign(3 * (size_t)nInt, 0.0); s.bH.assign(nInt, 0.0);
    if (nInt > 0) {
        std::vector<int> order;
        order.reserve(nInt);
        std::vector<int> st{0};
        while (!st.empty()) {  // reverse post-order via pre-order with children pushed after the node
            const int nd = st.back();
            st.pop_back();
            order.push_back(nd);
            if (s.left[nd] >= 0) st.push_back(s.left[nd]);
            if (s.right[nd] >= 0) st.push_back(s.right[nd]);
        }
        struct F32 { float mn[3], mx[3], c[3], m, H; };
        struct F64 { double mn[3], mx[3], c[3], m, H; };
        auto childF = [&](int ch, F32& f, F64& d) {
            if (ch < 0) {
                const int j = ~ch;
                const float p[3] = {(float)s.x[j], (float)s.y[j], (float)s.z[j]};
                const double pd[3] = {s.x[j], s.y[j], s.z[j]};
                for (int a = 0; a < 3; ++a) {
                    f.mn[a] = f.mx[a] = f.c[a] = p[a];
                    d.mn[a] = d.mx[a] = d.c[a] = pd[a];
                }
                f.m = (float)s.m[j];
                f.H = supportBoundF((float)s.h[j], (float)s.h0[j]);
                d.m = s.m[j];
                d.H = supportBoundD(s.h[j], s.h0[j]);
            } else {
                for (int a = 0; a < 3; ++a) {
                    f.mn[a] = s.mn4[4 * ch + a]; f.mx[a] = s.mx4[4 * ch + a]; f.c[a] = s.cm4[4 * ch + a];
                    d.mn[a] = s.bmn[3 * ch + a]; d.mx[a] = s.bmx[3 * ch + a];
                }
                d.c[0] = s.cmx[ch]; d.c[1] = s.cmy[ch]; d.c[2] = s.cmz[ch];
                f.m = s.cm4[4 * ch + 3]; f.H = s.mx4[4 * ch + 3];
                d.m = s.cmm[ch]; d.H = s.bH[ch];
            }
        };
        for (int oi = nInt - 1; oi >= 0; --oi) {
            const int nd = order[oi];
            F32 L, R; F64 LD, RD;
            childF(s.left[nd], L, LD);
            childF(s.right[nd], R, RD);
            float mn[3], mx[3], cc[3];
            for (int a = 0; a < 3; ++a) { mn[a] = std::fmin(L.mn[a], R.mn[a]); mx[a] = std::fmax(L.mx[a], R.mx[a]); }
            const float mS = L.m + R.m;
            if (mS > 0.0f) {
                const float invM = 1.0f / mS;
                for (int a = 0; a < 3; ++a) cc[a] = std::fmaf(L.m, L.c[a], R.m * R.c[a]) * invM;
            } else {
                for (int a = 0; a < 3; ++a) cc[a] = 0.5f * (L.c[a] + R.c[a]);
            }
            const float ext = std::fmax(mx[0] - mn[0], std::fmax(mx[1] - mn[1], mx[2] - mn[2]));
            for (int a = 0; a < 3; ++a) { s.cm4[4 * nd + a] = cc[a]; s.mn4[4 * nd + a] = mn[a]; s.mx4[4 * nd + a] = mx[a]; }
            s.cm4[4 * nd + 3] = mS; s.mn4[4 * nd + 3] = ext; s.mx4[4 * nd + 3] = std::fmax(L.H, R.H);
            const double mD = LD.m + RD.m;
            for (int a = 0; a < 3; ++a) {
                s.bmn[3 * nd + a] = std::min(LD.mn[a], RD.mn[a]);
                s.bmx[3 * nd + a] = std::max(LD.mx[a], RD.mx[a]);
            }
            if (mD > 0.0) {
                s.cmx[nd] = (LD.m * LD.c[0] + RD.m * RD.c[0]) / mD;
                s.cmy[nd] = (LD.m * LD.c[1] + RD.m * RD.c[1]) / mD;
                s.cmz[nd] = (LD.m * LD.c[2] + RD.m * RD.c[2]) / mD;
            } else {
                s.cmx[nd] = 0.5 * (LD.c[0] + RD.c[0]); s.cmy[nd] = 0.5 * (LD.c[1] + RD.c[1]); s.cmz[nd] = 0.5 * (LD.c[2] + RD.c[2]);
            }
            s.cmm[nd] = mD;
            s.bH[nd] = std::max(LD.H, RD.H);
        }
    }
    s.tree = true;
}

inline double boxDist2(const double* mn, const double* mx, double px, double py, double pz) {
    const double dx = std::max(0.0, std::max(mn[0] - px, px - mx[0]));
    const double dy = std::max(0.0, std::max(mn[1] - py, py - mx[1]));
    const double dz = std::max(0.0, std::max(mn[2] - pz, pz - mx[2]));
    return dx * dx + dy * dy + dz * dz;
}

// Visit every particle j != i with |x_i - x_j| possibly < max(Ri, bound_j) (conservative).
template <class F>
void forCandidates(const Seg& s, uint32_t i, double Ri, bool symmetric, F&& f) {
    if (s.n < 2) return;
    const double px = s.x[i], py = s.y[i], pz = s.z[i];
    int st[256];
    int sp = 0;
    st[sp++] = 0;
    std::vector<int> big;
    while (sp > 0 || !big.empty()) {
        int nd;
        if (!big.empty()) { nd = big.back(); big.pop_back(); } else nd = st[--sp];
        const int ch[2] = {s.left[nd], s.right[nd]};
        for (int k = 0; k < 2; ++k) {
            const int c = ch[k];
            if (c < 0) {
                const uint32_t j = (uint32_t)~c;
                if (j != i) f(j);
            } else {
                const double rad = (symmetric ? std::max(Ri, s.bH[c]) : Ri) * (1.0 + 1.0e-5) + 1.0e-12;
                if (boxDist2(&s.bmn[3 * c], &s.bmx[3 * c], px, py, pz) <= rad * rad) {
                    if (sp < 256) st[sp++] = c; else big.push_back(c);
                }
            }
        }
    }
}

// sphPass1Kernel: engine/sph_gravity.cu:645-788 (gather over r < H_i, then h update).
void pass1(or_ctx* c, Seg& s) {
    const uint32_t n = s.n;
    uint64_t pairs = 0;
    std::vector<double> hNew(n);
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : pairs)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const double hi = std::max(s.h[i], 1.0e-6), rhoI = std::max(s.rho[i], 1.0e-9);
        const double Hi = 2.0 * hi;
        double shep = (s.m[i] / rhoI) * (21.0 / (2.0 * kPi * Hi * Hi * Hi));
        int ngb = 1;
        double ns[3] = {0, 0, 0}, g[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        // FP32 decision replica: r2 < Hi^2 with the kernel's operation order
        const float xf = (float)s.x[i], yf = (float)s.y[i], zf = (float)s.z[i];
        const float Hif = 2.0f * std::fmax((float)s.h[i], 1.0e-6f);
        const float Hi2f = Hif * Hif;
        forCandidates(s, i, Hi, false, [&](uint32_t j) {
            const float r2f = dist2f(xf - (float)s.x[j], yf - (float)s.y[j], zf - (float)s.z[j]);
            if (!(r2f < Hi2f)) return;
            ++ngb;
            const double dx = s.x[i] - s.x[j], dy = s.y[i] - s.y[j], dz = s.z[i] - s.z[j];
            const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
            const double volJ = s.m[j] / std::max(s.rho[j], 1.0e-9);
            const double w = wendland(r, Hi);
            shep += volJ * w;
            ns[0] -= volJ * dx * w; ns[1] -= volJ * dy * w; ns[2] -= volJ * dz * w;
            const double gc = wendlandGrad(r, Hi);
            const double gg[3] = {gc * dx, gc * dy, gc * dz};
            const double dv[3] = {s.vx[j] - s.vx[i], s.vy[j] - s.vy[i], s.vz[j] - s.vz[i]};
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) g[a][b] += volJ * dv[a] * gg[b];
        });
        pairs += (uint64_t)(ngb - 1);
        s.shep[i] = shep;
        const double nn = std::sqrt(ns[0] * ns[0] + ns[1] * ns[1] + ns[2] * ns[2]);
        double f = std::min(1.0, std::max(0.0, 1.6 * std::max(1.0 - shep, nn / hi) - 0.15));
        if (ngb <= 2) f = 1.0;
        s.fs[i] = f;
        s.divv[i] = g[0][0] + g[1][1] + g[2][2];
        s.cu[0][i] = g[2][1] - g[1][2];
        s.cu[1][i] = g[0][2] - g[2][0];
        s.cu[2][i] = g[1][0] - g[0][1];
        s.ngb[i] = (uint32_t)ngb;
        // h' as the FP32 replica of the kernels' update (every later neighbour decision reads it):
        // the factor rounded once from FP64, then the product and the clamps in FP32
        const float hif = std::fmax((float)s.h[i], 1.0e-6f), h0f = std::fmax((float)s.h0[i], 1.0e-6f);
        const float facf = std::fmin(1.08f, std::fmax(0.92f, (float)std::cbrt(56.0 / (double)ngb)));
        hNew[i] = (double)std::fmin(6.0f * h0f, std::fmax(0.35f * h0f, hif * facf));
    }
    s.h = hNew;
    s.pairs1 = pairs;
}

void eosAll(or_ctx* c, Seg& s) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)s.n; ++i) {
        const EosOut o = eosEval(s.mat[i], s.rho[i], s.u[i], s.D[i], c->prm.gamma_hhe);
        s.P[i] = o.p; s.cs[i] = o.cs; s.T[i] = o.T; s.phase[i] = (uint8_t)o.phase;
    }
}

// gravityBarnesHutKernel: engine/sph_gravity.cu:810-961, with the opening decision shared by
// groups of G consecutive particles (the reference's G is the 32-lane warp, line 932).
void gravity(or_ctx* c, Seg& s) {
    const uint32_t n = s.n;
    const int G = c->G;
    const float th = std::fmax(0.0f, c->prm.theta);
    const float theta2 = th * th;
    const bool checkInside = theta2 > 0.333333f;
    const double soft = (double)c->prm.soft_factor;
    const float softF = c->prm.soft_factor;
    if (n == 1) {
        s.gx[0] = s.gy[0] = s.gz[0] = s.phi[0] = 0.0;
        s.ax[0] = s.ay[0] = s.az[0] = 0.0;
        s.interactions = 0; s.overflows = 0;
        return;
    }
    const uint32_t groups = (n + G - 1) / G;
    uint64_t inter = 0, ovf = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : inter, ovf)
    for (int64_t gi = 0; gi < (int64_t)groups; ++gi) {
        const uint32_t i0 = (uint32_t)gi * G, i1 = std::min<uint32_t>(i0 + G, n);
        const int cnt = (int)(i1 - i0);
        std::vector<double> ax(cnt, 0.0), ay(cnt, 0.0), az(cnt, 0.0), ph(cnt, 0.0), epsI(cnt);
        std::vector<float> xf(cnt), yf(cnt), zf(cnt);
        for (int l = 0; l < cnt; ++l) {
            const uint32_t i = i0 + l;
            xf[l] = (float)s.x[i]; yf[l] = (float)s.y[i]; zf[l] = (float)s.z[i];
            epsI[l] = soft * std::max(s.h0[i], 1.0e-6);
        }
        (void)softF;
        int st[ASTRAEA_STACK_DEPTH];
        int sp = 0;
        st[sp++] = 0;
        while (sp > 0) {
            const int nd = st[--sp];
            const int ch[2] = {s.left[nd], s.right[nd]};
            for (int k = 0; k < 2; ++k) {
                const int cc = ch[k];
                if (cc < 0) {
                    const uint32_t j = (uint32_t)~cc;
                    const double epsJ = soft * std::max(s.h0[j], 1.0e-6);
                    for (int l = 0; l < cnt; ++l) {
                        const uint32_t i = i0 + l;
                        if (i == j) continue;
                        const double dx = s.x[i] - s.x[j], dy = s.y[i] - s.y[j], dz = s.z[i] - s.z[j];
                        const double e = 0.5 * (epsI[l] + epsJ);
                        const double d2 = dx * dx + dy * dy + dz * dz + e * e + 1.0e-20;
                        const double ir = 1.0 / std::sqrt(d2), ir3 = ir * ir * ir;
                        const double f = -kG * s.m[j] * ir3;
                        ax[l] += f * dx; ay[l] += f * dy; az[l] += f * dz;
                        ph[l] -= kG * s.m[j] * ir;
                        ++inter;
                    }
                } else {
                    const float cx = s.cm4[4 * cc], cy = s.cm4[4 * cc + 1], cz = s.cm4[4 * cc + 2];
                    const float ext = s.mn4[4 * cc + 3];
                    bool anyOpen = false;
                    for (int l = 0; l < cnt && !anyOpen; ++l) {
                        const float r2 = dist2f(xf[l] - cx, yf[l] - cy, zf[l] - cz);
                        const float lhs = ext * ext, rhs = theta2 * r2;
                        bool acc = (lhs < rhs) && (r2 > 1.0e-20f);
                        if (acc && checkInside) {
                            const float* mn = &s.mn4[4 * cc];
                            const float* mx = &s.mx4[4 * cc];
                            if (xf[l] >= mn[0] && xf[l] <= mx[0] && yf[l] >= mn[1] && yf[l] <= mx[1] && zf[l] >= mn[2] &&
                                zf[l] <= mx[2])
                                acc = false;
                        }
                        if (!acc) anyOpen = true;
                    }
                    if (!anyOpen || sp >= ASTRAEA_STACK_DEPTH) {
                        if (anyOpen) ++ovf;
                        const double mN = s.cmm[cc];
                        for (int l = 0; l < cnt; ++l) {
                            const uint32_t i = i0 + l;
                            const double dx = s.x[i] - s.cmx[cc], dy = s.y[i] - s.cmy[cc], dz = s.z[i] - s.cmz[cc];
                            const double d2 = dx * dx + dy * dy + dz * dz + epsI[l] * epsI[l] + 1.0e-20;
                            const double ir = 1.0 / std::sqrt(d2), ir3 = ir * ir * ir;
                            const double f = -kG * mN * ir3;
                            ax[l] += f * dx; ay[l] += f * dy; az[l] += f * dz;
                            ph[l] -= kG * mN * ir;
                            ++inter;
                        }
                    } else {
                        st[sp++] = cc;
                    }
                }
            }
        }
        for (int l = 0; l < cnt; ++l) {
            const uint32_t i = i0 + l;
            s.gx[i] = ax[l]; s.gy[i] = ay[l]; s.gz[i] = az[l]; s.phi[i] = ph[l];
            s.ax[i] = ax[l]; s.ay[i] = ay[l]; s.az[i] = az[l];
        }
    }
    s.interactions = inter;
    s.overflows = ovf;
}

// effectiveSphPressure: engine/sph_gravity.cu:965-969
inline double effP(double p, int phase) { return phase != kSolid ? std::max(p, 0.0) : p; }

// sphPass2Kernel: engine/sph_gravity.cu:971-1287
void pass2(or_ctx* c, Seg& s) {
    const uint32_t n = s.n;
    uint64_t pairs = 0;
    std::vector<double> Hs(n), pE(n), csF(n), fB(n);
    std::vector<float> Hf(n), xf(n), yf(n), zf(n);  // FP32 replicas for the pair decision
    for (uint32_t i = 0; i < n; ++i) {
        const double hi = std::max(s.h[i], 1.0e-6);
        Hs[i] = 2.0 * hi;
        Hf[i] = 2.0f * std::fmax((float)s.h[i], 1.0e-6f);
        xf[i] = (float)s.x[i]; yf[i] = (float)s.y[i]; zf[i] = (float)s.z[i];
        pE[i] = effP(s.P[i], s.phase[i]);
        csF[i] = std::max(s.cs[i], 0.05);
        fB[i] = balsara(s.divv[i], s.cu[0][i], s.cu[1][i], s.cu[2][i], csF[i], hi);
    }
    const double coolMin = (double)c->prm.cooling_min_temp_k;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : pairs)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const double hi = std::max(s.h[i], 1.0e-6), Hi = Hs[i];
        const double rhoI = std::max(s.rho[i], 1.0e-9), iR2 = 1.0 / (rhoI * rhoI);
        const double pI = pE[i], csI = csF[i], tI = s.T[i], dmgI = s.D[i], fsI = s.fs[i];
        const int matI = s.mat[i];
        const double* sI[6];
        double si[6];
        for (int a = 0; a < 6; ++a) { si[a] = s.s[a][i]; sI[a] = &si[a]; }
        (void)sI;
        const double fbI = fB[i], Ri = pI < 0.0 ? 0.3 * (-pI) * iR2 : 0.0, pr2I = pI * iR2;
        double a[3] = {0, 0, 0}, dC = 0, dDf = 0, duP = 0, duAv = 0, duT = 0, duS = 0;
        double g[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        forCandidates(s, i, Hi, true, [&](uint32_t j) {
            // FP32 replica of the kernels' decision r^2 < max(H_i, H_j)^2
            const float r2f = dist2f(xf[i] - xf[j], yf[i] - yf[j], zf[i] - zf[j]);
            const float Hmf = std::fmax(Hf[i], Hf[j]);
            if (!(r2f < Hmf * Hmf)) return;
            const double dx = s.x[i] - s.x[j], dy = s.y[i] - s.y[j], dz = s.z[i] - s.z[j];
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double Hj = Hs[j];
            ++pairs;
            const double hj = 0.5 * Hj;
            const double mj = s.m[j], rhoJ = std::max(s.rho[j], 1.0e-9), pJ = pE[j], csJ = csF[j];
            const double dv[3] = {s.vx[i] - s.vx[j], s.vy[i] - s.vy[j], s.vz[i] - s.vz[j]};
            const double vdr = dv[0] * dx + dv[1] * dy + dv[2] * dz;
            const double r = std::sqrt(r2);
            const double gI = wendlandGrad(r, Hi), gJ = wendlandGrad(r, Hj), gb = 0.5 * (gI + gJ);
            const double gv[3] = {gb * dx, gb * dy, gb * dz};
            const double iRj2 = 1.0 / (rhoJ * rhoJ);
            const double vdg = dv[0] * gv[0] + dv[1] * gv[1] + dv[2] * gv[2];
            dC += (mj / rhoJ) * vdg;
            if (s.mat[j] == matI) {
                const double rdg = -(dx * gv[0] + dy * gv[1] + dz * gv[2]);
                dDf += (mj / rhoJ) * (2.0 * (rhoJ - rhoI) / (r2 + 0.01 * Hi * Hi)) * rdg;
            }
            double pi = 0.0;
            if (vdr < 0.0) {
                const double hb = 0.5 * (hi + hj), cb = 0.5 * (csI + csJ), rb = 0.5 * (rhoI + rhoJ);
                const double mu = hb * vdr / (r2 + 0.01 * hb * hb);
                pi = ((-cb * mu + 2.0 * mu * mu) / rb) * (0.5 * (fbI + fB[j]));
            }
            double tens = 0.0;
            const double Rj = pJ < 0.0 ? 0.3 * (-pJ) * iRj2 : 0.0;
            if (Ri + Rj > 0.0) {
                const double Hb = 0.5 * (Hi + Hj);
                if (r < Hb) {
                    const double q = r / Hb, o = 1.0 - q, o2 = o * o;
                    const double fr = (o2 * o2 * (1.0 + 4.0 * q)) / 0.6328125;
                    const double fr2 = fr * fr;
                    tens = (Ri + Rj) * fr2 * fr2;
                }
            }
            const double iso = -mj * (pr2I + pJ * iRj2 + pi + tens);
            for (int q = 0; q < 3; ++q) a[q] += iso * gv[q];
            duP += mj * vdg;
            duAv += 0.5 * mj * pi * vdg;
            duT += 0.5 * mj * tens * vdg;
            double st[6];
            for (int q = 0; q < 6; ++q) st[q] = si[q] * iR2 + s.s[q][j] * iRj2;
            // tensor order: xx yy zz xy xz yz
            a[0] += mj * (st[0] * gv[0] + st[3] * gv[1] + st[4] * gv[2]);
            a[1] += mj * (st[3] * gv[0] + st[1] * gv[1] + st[5] * gv[2]);
            a[2] += mj * (st[4] * gv[0] + st[5] * gv[1] + st[2] * gv[2]);
            const double w[3] = {-dv[0], -dv[1], -dv[2]};
            duS += mj * ((si[0] * w[0] + si[3] * w[1] + si[4] * w[2]) * gv[0] +
                         (si[3] * w[0] + si[1] * w[1] + si[5] * w[2]) * gv[1] +
                         (si[4] * w[0] + si[5] * w[1] + si[2] * w[2]) * gv[2]);
            const double volJ = mj / rhoJ;
            const double gi[3] = {gI * dx, gI * dy, gI * dz};
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) g[p][q] += volJ * w[p] * gi[q];
        });
        s.ax[i] += a[0]; s.ay[i] += a[1]; s.az[i] += a[2];
        const double mask = s.shep[i] > 0.85 ? (1.0 - fsI) : 0.0;
        s.drho[i] = rhoI * dC + 0.05 * Hi * csI * mask * dDf;
        // radiative cooling (FP64 in the reference too): sph_gravity.cu:1219-1235
        const double tAmb = (s.living && s.body[i] == 0 && (float)dmgI < 0.8f) ? 255.0 : 3.0;
        double pRad = 0.0;
        if (tI > coolMin) {
            const double mSi = std::max(s.m[i] * kMUnit, 1.0), rhoSi = std::max(rhoI * kRhoUnit, 1.0e-3);
            const double rp = std::cbrt((3.0 * mSi) / (4.0 * kPi * rhoSi));
            const double area = kPi * rp * rp;
            const double dT4 = std::max(0.0, tI * tI * tI * tI - tAmb * tAmb * tAmb * tAmb);
            pRad = fsI * 0.9 * kSigmaSB * area * dT4;
        }
        s.prad[i] = pRad;
        const double duRad = -pRad / (std::max(s.m[i], 1.0e-30) * 1.0e27);
        s.du[i] = pr2I * duP + duAv + duT + iR2 * duS + duRad;
        // Jaumann stress rate and damage: sph_gravity.cu:1237-1286
        if (!strengthOn(c, matI)) {
            for (int q = 0; q < 6; ++q) s.dS[q][i] = 0.0;
            s.dD[i] = 0.0;
            continue;
        }
        const Mat& mp = matOf(matI);
        const double muS = mp.mu / kPUnit, tM = mp.tMelt;
        const double tn = tI / tM, soft = std::max(0.0, 1.0 - tn * tn);
        const double muE = (1.0 - dmgI) * muS * soft;
        const double tr = (g[0][0] + g[1][1] + g[2][2]) / 3.0;
        const double exx = g[0][0] - tr, eyy = g[1][1] - tr, ezz = g[2][2] - tr;
        const double exy = 0.5 * (g[0][1] + g[1][0]), exz = 0.5 * (g[0][2] + g[2][0]), eyz = 0.5 * (g[1][2] + g[2][1]);
        const double rxy = 0.5 * (g[0][1] - g[1][0]), rxz = 0.5 * (g[0][2] - g[2][0]), ryz = 0.5 * (g[1][2] - g[2][1]);
        const double ryx = -rxy, rzx = -rxz, rzy = -ryz;
        const double sxx = si[0], syy = si[1], szz = si[2], sxy = si[3], sxz = si[4], syz = si[5];
        s.dS[0][i] = 2.0 * muE * exx + 2.0 * (sxy * rxy + sxz * rxz);
        s.dS[1][i] = 2.0 * muE * eyy + 2.0 * (sxy * ryx + syz * ryz);
        s.dS[2][i] = 2.0 * muE * ezz + 2.0 * (sxz * rzx + syz * rzy);
        s.dS[3][i] = 2.0 * muE * exy + sxx * ryx + sxz * ryz + syy * rxy + syz * rxz;
        s.dS[4][i] = 2.0 * muE * exz + sxx * rzx + sxy * rzy + szz * rxz + syz * rxy;
        s.dS[5][i] = 2.0 * muE * eyz + sxy * rzx + syy * rzy + sxz * ryx + szz * ryz;
        if (tI < tM && dmgI < 1.0) {
            const double sm = maxPrincipal(pI, si);
            const double K = mp.bulk / kPUnit;
            const double stiff = std::max(K + (4.0 / 3.0) * muS, 1.0e-12);
            const double et = std::max(0.0, sm) / stiff;
            const double ef = std::max(s.eps[i], 1.0e-8);
            const double cc = std::max(csI, std::sqrt((K + (4.0 / 3.0) * muS) / rhoI));
            s.dD[i] = et > ef ? 0.4 * cc / hi : 0.0;
        } else {
            s.dD[i] = 0.0;
        }
    }
    s.pairs2 = pairs;
}

// leapfrogKickHalfKernel: engine/sph_gravity.cu:1291-1374
void kick(or_ctx* c, Seg& s, double hdt, bool clampThermal) {
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)s.n; ++ii) {
        const size_t i = (size_t)ii;
        s.vx[i] += hdt * s.ax[i]; s.vy[i] += hdt * s.ay[i]; s.vz[i] += hdt * s.az[i];
        const double rhoN = std::max(s.rho[i] + hdt * s.drho[i], 1.0e-9);
        s.rho[i] = rhoN;
        const int mi = matIdx(s.mat[i]);
        const Mat& mp = kMat[mi];
        double uN = s.u[i] + hdt * s.du[i];
        if (clampThermal) uN = std::max(uN, coldEnergy(mi, rhoN) + 3.0 * mp.cv / kUUnit);
        else uN = std::max(uN, 0.0);
        s.u[i] = uN;
        double dc = s.Dc[i] + hdt * s.dD[i];
        dc = std::min(1.0, std::max(s.Dc[i], dc));
        const double D = dc * dc * dc;
        s.Dc[i] = dc; s.D[i] = D;
        if (clampThermal) {
            const EosOut o = eosEval(mi, rhoN, uN, D, c->prm.gamma_hhe);
            s.P[i] = o.p; s.cs[i] = o.cs; s.T[i] = o.T; s.phase[i] = (uint8_t)o.phase;
        }
        if (!strengthOn(c, mi) || s.T[i] >= mp.tMelt) {
            for (int q = 0; q < 6; ++q) s.s[q][i] = 0.0;
            continue;
        }
        double t[6];
        for (int q = 0; q < 6; ++q) t[q] = s.s[q][i] + hdt * s.dS[q][i];
        const double tr = (t[0] + t[1] + t[2]) / 3.0;
        t[0] -= tr; t[1] -= tr; t[2] -= tr;
        const double y0 = mp.Y0 / kPUnit;
        const double yI = (1.0 - D) * y0 * std::max(0.0, 1.0 - s.T[i] / mp.tMelt);
        const double yF = D * std::min(y0, 0.6 * std::max(0.0, s.P[i]));
        const double Y = std::max(yI, yF);
        const double j2 = 0.5 * (t[0] * t[0] + t[1] * t[1] + t[2] * t[2] + 2.0 * (t[3] * t[3] + t[4] * t[4] + t[5] * t[5]));
        const double eq = std::sqrt(3.0 * j2 + 1.0e-30);
        const double sc = Y <= 0.0 ? 0.0 : std::min(1.0, Y / eq);
        for (int q = 0; q < 6; ++q) s.s[q][i] = t[q] * sc;
    }
}

// leapfrogDriftKernel: engine/sph_gravity.cu:1376-1489 (XSPH 0.25 over r < max(H_i, H_j),
// positions before the drift, neighbour search on the tree of the previous sort).
void drift(or_ctx* c, Seg& s, double dt) {
    (void)c;
    const uint32_t n = s.n;
    std::vector<double> nx(n), ny(n), nz(n);
    std::vector<double> Hs(n);
    std::vector<float> Hf(n), xf(n), yf(n), zf(n);  // FP32 replicas for the pair decision
    for (uint32_t i = 0; i < n; ++i) {
        Hs[i] = 2.0 * std::max(s.h[i], 1.0e-6);
        Hf[i] = 2.0f * std::fmax((float)s.h[i], 1.0e-6f);
        xf[i] = (float)s.x[i]; yf[i] = (float)s.y[i]; zf[i] = (float)s.z[i];
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const double Hi = Hs[i], rhoI = std::max(s.rho[i], 1.0e-9);
        double xs[3] = {0, 0, 0};
        if (s.tree && n >= 2) {
            forCandidates(s, i, Hi, true, [&](uint32_t j) {
                const float r2f = dist2f(xf[i] - xf[j], yf[i] - yf[j], zf[i] - zf[j]);
                const float Hmf = std::fmax(Hf[i], Hf[j]);
                if (!(r2f < Hmf * Hmf)) return;
                const double dx = s.x[i] - s.x[j], dy = s.y[i] - s.y[j], dz = s.z[i] - s.z[j];
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double r = std::sqrt(r2);
                const double wb = 0.5 * (wendland(r, Hi) + wendland(r, Hs[j]));
                const double rb = 0.5 * (rhoI + std::max(s.rho[j], 1.0e-9));
                const double co = (s.m[j] / rb) * wb;
                xs[0] += co * (s.vx[j] - s.vx[i]); xs[1] += co * (s.vy[j] - s.vy[i]); xs[2] += co * (s.vz[j] - s.vz[i]);
            });
        }
        nx[i] = s.x[i] + dt * (s.vx[i] + 0.25 * xs[0]);
        ny[i] = s.y[i] + dt * (s.vy[i] + 0.25 * xs[1]);
        nz[i] = s.z[i] + dt * (s.vz[i] + 0.25 * xs[2]);
    }
    s.x = nx; s.y = ny; s.z = nz;
}

// reduceDtMinKernel + launchReduceDtPsu: engine/sph_gravity.cu:1493-1550, 2136-2161
double reduceDt(or_ctx* c, Seg& s, double maxDt) {
    const double floorDt = std::min(1.0e-5, maxDt);
    const double upper = s.prevDt > 0.0 ? std::min(s.prevDt * 1.15, maxDt) : maxDt;
    if (s.n == 0) return upper;
    double mn = 1.0e30;
    for (uint32_t i = 0; i < s.n; ++i) {
        const double hi = std::max(s.h[i], 1.0e-6), cs = std::max(s.cs[i], 0.05);
        const double v = std::sqrt(s.vx[i] * s.vx[i] + s.vy[i] * s.vy[i] + s.vz[i] * s.vz[i]);
        const double dC = (double)c->prm.cfl_factor * hi / (2.2 * cs + v + 1.0e-6);
        const double a = std::sqrt(s.ax[i] * s.ax[i] + s.ay[i] * s.ay[i] + s.az[i] * s.az[i]);
        const double dA = 0.25 * std::sqrt(hi / (a + 1.0e-8));
        const int mi = matIdx(s.mat[i]);
        const double u0 = (double)(float)((kMat[mi].cv * 300.0) / kUUnit);
        const double uu = std::max(s.u[i], 0.0), us = std::max(uu, u0);
        const double dT = 0.20 * (uu + 0.01 * us) / (std::abs(s.du[i]) + 1.0e-8);
        const double d = std::min(dC, std::min(dA, dT));
        if (std::isfinite(d) && d > 0.0) mn = std::min(mn, d);
    }
    return std::min(upper, std::max(floorDt, mn));
}

double sumRad(const Seg& s) {
    double t = 0.0;
    for (uint32_t i = 0; i < s.n; ++i) t += s.prad[i];
    return t;
}

Seg* segOf(or_ctx* c, uint32_t k) { return k < c->seg.size() ? &c->seg[k] : nullptr; }

// Generic field accessors
std::vector<double>* f64Field(Seg& s, int f) {
    switch (f) {
        case ASTRAEA_F_POS_X: case ASTRAEA_F_POS64_X: return &s.x;
        case ASTRAEA_F_POS_Y: case ASTRAEA_F_POS64_Y: return &s.y;
        case ASTRAEA_F_POS_Z: case ASTRAEA_F_POS64_Z: return &s.z;
        case ASTRAEA_F_VEL_X: case ASTRAEA_F_VEL64_X: return &s.vx;
        case ASTRAEA_F_VEL_Y: case ASTRAEA_F_VEL64_Y: return &s.vy;
        case ASTRAEA_F_VEL_Z: case ASTRAEA_F_VEL64_Z: return &s.vz;
        case ASTRAEA_F_MASS: return &s.m;
        case ASTRAEA_F_H: return &s.h;
        case ASTRAEA_F_H0: return &s.h0;
        case ASTRAEA_F_RHO: return &s.rho;
        case ASTRAEA_F_U: return &s.u;
        case ASTRAEA_F_SXX: return &s.s[0];
        case ASTRAEA_F_SYY: return &s.s[1];
        case ASTRAEA_F_SZZ: return &s.s[2];
        case ASTRAEA_F_SXY: return &s.s[3];
        case ASTRAEA_F_SXZ: return &s.s[4];
        case ASTRAEA_F_SYZ: return &s.s[5];
        case ASTRAEA_F_DAMAGE: return &s.D;
        case ASTRAEA_F_DAMAGE_CBRT: return &s.Dc;
        case ASTRAEA_F_EPS_FAIL: return &s.eps;
        case ASTRAEA_F_ACC_X: return &s.ax;
        case ASTRAEA_F_ACC_Y: return &s.ay;
        case ASTRAEA_F_ACC_Z: return &s.az;
        case ASTRAEA_F_ACC_GRAV_X: return &s.gx;
        case ASTRAEA_F_ACC_GRAV_Y: return &s.gy;
        case ASTRAEA_F_ACC_GRAV_Z: return &s.gz;
        case ASTRAEA_F_PHI: return &s.phi;
        case ASTRAEA_F_DRHO_DT: return &s.drho;
        case ASTRAEA_F_DU_DT: return &s.du;
        case ASTRAEA_F_PRESSURE: return &s.P;
        case ASTRAEA_F_SOUND_SPEED: return &s.cs;
        case ASTRAEA_F_TEMP_K: return &s.T;
        case ASTRAEA_F_SHEPARD: return &s.shep;
        case ASTRAEA_F_SURF_FACTOR: return &s.fs;
        case ASTRAEA_F_DIV_V: return &s.divv;
        case ASTRAEA_F_CURL_X: return &s.cu[0];
        case ASTRAEA_F_CURL_Y: return &s.cu[1];
        case ASTRAEA_F_CURL_Z: return &s.cu[2];
        case ASTRAEA_F_DS_XX: return &s.dS[0];
        case ASTRAEA_F_DS_YY: return &s.dS[1];
        case ASTRAEA_F_DS_ZZ: return &s.dS[2];
        case ASTRAEA_F_DS_XY: return &s.dS[3];
        case ASTRAEA_F_DS_XZ: return &s.dS[4];
        case ASTRAEA_F_DS_YZ: return &s.dS[5];
        case ASTRAEA_F_DDAMAGE_CBRT_DT: return &s.dD;
        case ASTRAEA_F_RAD_POWER_W: return &s.prad;
        default: return nullptr;
    }
}
std::vector<uint8_t>* u8Field(Seg& s, int f) {
    switch (f) {
        case ASTRAEA_F_MATERIAL: return &s.mat;
        case ASTRAEA_F_BODY: return &s.body;
        case ASTRAEA_F_TRACER: return &s.tracer;
        case ASTRAEA_F_PHASE: return &s.phase;
        default: return nullptr;
    }
}
bool isF64Field(int f) { return f >= ASTRAEA_F_POS64_X && f <= ASTRAEA_F_VEL64_Z; }

template <class F>
int forSegs(or_ctx* c, F&& f) {
    if (!c) return ASTRAEA_E_INVALID;
    for (auto& s : c->seg) f(s);
    return ASTRAEA_OK;
}

}  // namespace

extern "C" {

or_ctx* or_create(void) {
    auto* c = new or_ctx();
    astraea_params p{};
    p.theta = 0.5f; p.soft_factor = 0.8f; p.cfl_factor = 0.25f; p.gamma_hhe = 1.4f;
    p.cooling_min_temp_k = 0.0f; p.strength_mask = 0x17u; p.precision = 0u;
    c->prm = p;
    return c;
}
void or_destroy(or_ctx* c) { delete c; }

int or_set_params(or_ctx* c, const astraea_params* p, int group_size) {
    if (!c || !p || group_size < 1) return ASTRAEA_E_INVALID;
    c->prm = *p;
    c->G = group_size;
    return ASTRAEA_OK;
}

int or_set_layout(or_ctx* c, uint32_t ns, const uint32_t* counts) {
    if (!c || (ns > 0 && !counts)) return ASTRAEA_E_INVALID;
    c->seg.assign(ns, Seg());
    uint32_t base = 0;
    for (uint32_t k = 0; k < ns; ++k) {
        c->seg[k].resize(counts[k]);
        for (uint32_t i = 0; i < counts[k]; ++i) c->seg[k].oid[i] = base + i;
        base += counts[k];
    }
    return ASTRAEA_OK;
}

int or_upload_field(or_ctx* c, int f, const void* src) {
    if (!c || !src) return ASTRAEA_E_INVALID;
    size_t off = 0;
    for (auto& s : c->seg) {
        if (auto* v = f64Field(s, f)) {
            if (isF64Field(f)) for (uint32_t i = 0; i < s.n; ++i) (*v)[i] = ((const double*)src)[off + i];
            else for (uint32_t i = 0; i < s.n; ++i) (*v)[i] = (double)((const float*)src)[off + i];
        } else if (auto* b = u8Field(s, f)) {
            for (uint32_t i = 0; i < s.n; ++i) (*b)[i] = ((const uint8_t*)src)[off + i];
        } else if (f == ASTRAEA_F_ORIGINAL_ID) {
            for (uint32_t i = 0; i < s.n; ++i) s.oid[i] = ((const uint32_t*)src)[off + i];
        } else {
            return ASTRAEA_E_INVALID;
        }
        // positions, masses and smoothing lengths shape the tree; other fields leave it valid
        if ((f >= ASTRAEA_F_POS_X && f <= ASTRAEA_F_POS_Z) || f == ASTRAEA_F_MASS || f == ASTRAEA_F_H ||
            f == ASTRAEA_F_H0 || (f >= ASTRAEA_F_POS64_X && f <= ASTRAEA_F_POS64_Z))
            s.tree = false;
        off += s.n;
    }
    return ASTRAEA_OK;
}

int or_download_field(or_ctx* c, int f, void* dst) {
    if (!c || !dst) return ASTRAEA_E_INVALID;
    size_t off = 0;
    for (auto& s : c->seg) {
        if (auto* v = f64Field(s, f)) {
            if (isF64Field(f)) for (uint32_t i = 0; i < s.n; ++i) ((double*)dst)[off + i] = (*v)[i];
            else for (uint32_t i = 0; i < s.n; ++i) ((float*)dst)[off + i] = (float)(*v)[i];
        } else if (auto* b = u8Field(s, f)) {
            for (uint32_t i = 0; i < s.n; ++i) ((uint8_t*)dst)[off + i] = (*b)[i];
        } else if (f == ASTRAEA_F_ORIGINAL_ID) {
            for (uint32_t i = 0; i < s.n; ++i) ((uint32_t*)dst)[off + i] = s.oid[i];
        } else if (f == ASTRAEA_F_NGB_COUNT) {
            for (uint32_t i = 0; i < s.n; ++i) ((uint32_t*)dst)[off + i] = s.ngb[i];
        } else {
            return ASTRAEA_E_INVALID;
        }
        off += s.n;
    }
    return ASTRAEA_OK;
}

int or_download_field_f64(or_ctx* c, int f, double* dst) {
    if (!c || !dst) return ASTRAEA_E_INVALID;
    size_t off = 0;
    for (auto& s : c->seg) {
        auto* v = f64Field(s, f);
        if (!v) return ASTRAEA_E_INVALID;
        for (uint32_t i = 0; i < s.n; ++i) dst[off + i] = (*v)[i];
        off += s.n;
    }
    return ASTRAEA_OK;
}

// launchClearStats / clearStatsKernel: engine/sph_gravity.cu:624-641
int or_clear_stats(or_ctx* c) {
    return forSegs(c, [](Seg& s) {
        for (auto* v : {&s.ax, &s.ay, &s.az, &s.gx, &s.gy, &s.gz, &s.drho, &s.du, &s.prad, &s.dS[0], &s.dS[1], &s.dS[2],
                        &s.dS[3], &s.dS[4], &s.dS[5], &s.dD, &s.divv, &s.cu[0], &s.cu[1], &s.cu[2], &s.fs, &s.phi})
            std::fill(v->begin(), v->end(), 0.0);
        std::fill(s.shep.begin(), s.shep.end(), 1.0);
    });
}
int or_sort_and_build(or_ctx* c) { return forSegs(c, [c](Seg& s) { sortAndBuild(c, s); }); }
int or_pass1(or_ctx* c) { return forSegs(c, [c](Seg& s) { if (s.tree) pass1(c, s); }); }
int or_eos(or_ctx* c) { return forSegs(c, [c](Seg& s) { eosAll(c, s); }); }
int or_gravity(or_ctx* c) { return forSegs(c, [c](Seg& s) { if (s.tree) gravity(c, s); }); }
int or_pass2(or_ctx* c) { return forSegs(c, [c](Seg& s) { if (s.tree) pass2(c, s); }); }

int or_reduce_dt(or_ctx* c, double maxDt, double* out) {
    if (!c) return ASTRAEA_E_INVALID;
    for (size_t k = 0; k < c->seg.size(); ++k) {
        Seg& s = c->seg[k];
        s.dt = reduceDt(c, s, maxDt);
        s.prevDt = s.dt;
        if (out) out[k] = s.dt;
    }
    return ASTRAEA_OK;
}
int or_set_dt(or_ctx* c, const double* dt) {
    if (!c || !dt) return ASTRAEA_E_INVALID;
    for (size_t k = 0; k < c->seg.size(); ++k) { c->seg[k].dt = dt[k]; c->seg[k].prevDt = dt[k]; }
    return ASTRAEA_OK;
}
int or_get_dt(or_ctx* c, double* dt) {
    if (!c || !dt) return ASTRAEA_E_INVALID;
    for (size_t k = 0; k < c->seg.size(); ++k) dt[k] = c->seg[k].dt;
    return ASTRAEA_OK;
}
int or_kick_half(or_ctx* c, int clamp) { return forSegs(c, [c, clamp](Seg& s) { kick(c, s, 0.5 * s.dt, clamp != 0); }); }
int or_drift(or_ctx* c) { return forSegs(c, [c](Seg& s) { drift(c, s, s.dt); }); }

// evaluateInitialForcesOnDevice: engine/simulation.cpp:525-536
int or_initial_forces(or_ctx* c) {
    if (!c) return ASTRAEA_E_INVALID;
    or_clear_stats(c);
    or_sort_and_build(c);
    or_pass1(c);
    or_eos(c);
    or_gravity(c);
    or_pass2(c);
    for (auto& s : c->seg) s.prevRadW = sumRad(s);
    return ASTRAEA_OK;
}

// Simulation::step kernel sequence: engine/simulation.cpp:1711-1753
int or_step(or_ctx* c, double maxDt, uint32_t steps) {
    if (!c) return ASTRAEA_E_INVALID;
    for (uint32_t st = 0; st < steps; ++st) {
        for (auto& s : c->seg) {
            if (s.n == 0) continue;
            s.dt = reduceDt(c, s, maxDt);
            s.prevDt = s.dt;
            kick(c, s, 0.5 * s.dt, false);
            drift(c, s, s.dt);
            sortAndBuild(c, s);
            pass1(c, s);
            eosAll(c, s);
            gravity(c, s);
            pass2(c, s);
            kick(c, s, 0.5 * s.dt, false);
            eosAll(c, s);
            const double dtSec = s.dt * 1.0e3;
            const double cur = sumRad(s);
            s.erad += 0.5 * (s.prevRadW + cur) * dtSec / 1.0e30;
            s.prevRadW = cur;
            s.simTime += s.dt;
        }
    }
    return ASTRAEA_OK;
}

// One relaxation KDK step: engine/simulation.cpp:603-615 (dt_max 0.02, thermal clamp in both kicks).
static void relaxKdk(or_ctx* c, Seg& s) {
    s.dt = reduceDt(c, s, 0.02);
    s.prevDt = s.dt;
    kick(c, s, 0.5 * s.dt, true);
    drift(c, s, s.dt);
    sortAndBuild(c, s);
    pass1(c, s);
    eosAll(c, s);
    gravity(c, s);
    pass2(c, s);
    kick(c, s, 0.5 * s.dt, true);
}

// u from the anchor temperature, stress and damage cleared: engine/simulation.cpp:669-700.
static void reanchor(Seg& s, const std::vector<double>& anchor) {
    for (uint32_t i = 0; i < s.n; ++i) {
        const int mi = matIdx(s.mat[i]);
        const double t = s.oid[i] < anchor.size() ? anchor[s.oid[i]] : 288.0;
        s.u[i] = coldEnergy(mi, s.rho[i]) + kMat[mi].cv / kUUnit * t;
        for (int q = 0; q < 6; ++q) { s.s[q][i] = 0.0; s.dS[q][i] = 0.0; }
        s.D[i] = 0.0; s.Dc[i] = 0.0; s.dD[i] = 0.0;
    }
}

// Simulation::relaxLoadedSingleBody, engine/simulation.cpp:585-781, one body per segment.
int or_relax(or_ctx* c, uint32_t steps, const double* anchorTemp, uint32_t nAnchor, const double* nominalRadius,
             double* out) {
    if (!c || !nominalRadius || !out) return ASTRAEA_E_INVALID;
    or_initial_forces(c);
    std::vector<double> anchor;
    if (anchorTemp) {
        anchor.assign(anchorTemp, anchorTemp + nAnchor);
    } else {
        anchor.assign(nAnchor, 288.0);
        for (auto& s : c->seg)
            for (uint32_t i = 0; i < s.n; ++i)
                if (s.oid[i] < nAnchor) anchor[s.oid[i]] = s.T[i];
    }
    for (size_t k = 0; k < c->seg.size(); ++k) {
        Seg& s = c->seg[k];
        s.prevDt = 0.0;
        for (uint32_t st = 0; st < steps && s.n > 0; ++st) {
            relaxKdk(c, s);
            double M = 0.0, cx[3] = {}, cv[3] = {};
            for (uint32_t i = 0; i < s.n; ++i) {
                M += s.m[i];
                cx[0] += s.m[i] * s.x[i]; cx[1] += s.m[i] * s.y[i]; cx[2] += s.m[i] * s.z[i];
                cv[0] += s.m[i] * s.vx[i]; cv[1] += s.m[i] * s.vy[i]; cv[2] += s.m[i] * s.vz[i];
            }
            if (M > 0.0)
                for (int a = 0; a < 3; ++a) { cx[a] /= M; cv[a] /= M; }
            for (uint32_t i = 0; i < s.n; ++i) {
                s.x[i] -= cx[0]; s.y[i] -= cx[1]; s.z[i] -= cx[2];
                s.vx[i] = 0.85 * (s.vx[i] - cv[0]);
                s.vy[i] = 0.85 * (s.vy[i] - cv[1]);
                s.vz[i] = 0.85 * (s.vz[i] - cv[2]);
            }
            reanchor(s, anchor);
            if (st + 1 < steps) {
                eosAll(c, s);
                gravity(c, s);
                pass2(c, s);
            }
        }
        if (steps > 0 && s.n > 0) {
            sortAndBuild(c, s);
            pass1(c, s);
            eosAll(c, s);
            gravity(c, s);
            pass2(c, s);
            for (int d = 0; d < 4; ++d) relaxKdk(c, s);
            eosAll(c, s);
        }
        double M = 0.0, mv2 = 0.0, rMax = nominalRadius[k];
        for (uint32_t i = 0; i < s.n; ++i) {
            M += s.m[i];
            mv2 += s.m[i] * (s.vx[i] * s.vx[i] + s.vy[i] * s.vy[i] + s.vz[i] * s.vz[i]);
            rMax = std::max(rMax, std::sqrt(s.x[i] * s.x[i] + s.y[i] * s.y[i] + s.z[i] * s.z[i]));
        }
        const double vRms = M > 0.0 ? std::sqrt(mv2 / M) : 0.0;
        const double vEsc = std::sqrt(2.0 * kG * M / std::max(rMax, 1.0e-3));
        out[k] = vEsc > 0.0 ? vRms / vEsc : 0.0;
        if (steps > 0 && s.n > 0) {
            reanchor(s, anchor);
            eosAll(c, s);
        }
    }
    return ASTRAEA_OK;
}

// Conservation reductions: engine/simulation.cpp:985-1087
int or_energy(or_ctx* c, double* out) {
    if (!c || !out) return ASTRAEA_E_INVALID;
    for (size_t k = 0; k < c->seg.size(); ++k) {
        Seg& s = c->seg[k];
        double* o = out + k * ASTRAEA_ENERGY_STRIDE;
        for (int q = 0; q < ASTRAEA_ENERGY_STRIDE; ++q) o[q] = 0.0;
        for (uint32_t i = 0; i < s.n; ++i) {
            const double m = s.m[i], x = s.x[i], y = s.y[i], z = s.z[i], vx = s.vx[i], vy = s.vy[i], vz = s.vz[i];
            o[ASTRAEA_EN_MASS] += m;
            o[ASTRAEA_EN_KINETIC] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
            o[ASTRAEA_EN_POTENTIAL] += 0.5 * m * s.phi[i];
            o[ASTRAEA_EN_INTERNAL] += m * s.u[i];
            o[ASTRAEA_EN_MOM_X] += m * vx; o[ASTRAEA_EN_MOM_Y] += m * vy; o[ASTRAEA_EN_MOM_Z] += m * vz;
            o[ASTRAEA_EN_ANG_X] += m * (y * vz - z * vy);
            o[ASTRAEA_EN_ANG_Y] += m * (z * vx - x * vz);
            o[ASTRAEA_EN_ANG_Z] += m * (x * vy - y * vx);
            o[ASTRAEA_EN_RAD_POWER_W] += s.prad[i];
            o[ASTRAEA_EN_MAX_TEMP] = std::max(o[ASTRAEA_EN_MAX_TEMP], s.T[i]);
            o[ASTRAEA_EN_MAX_PRESSURE] = std::max(o[ASTRAEA_EN_MAX_PRESSURE], s.P[i]);
            o[ASTRAEA_EN_COM_X] += m * x; o[ASTRAEA_EN_COM_Y] += m * y; o[ASTRAEA_EN_COM_Z] += m * z;
        }
        if (o[ASTRAEA_EN_MASS] > 0.0)
            for (int q = ASTRAEA_EN_COM_X; q <= ASTRAEA_EN_COM_Z; ++q) o[q] /= o[ASTRAEA_EN_MASS];
        o[ASTRAEA_EN_RADIATED] = s.erad;
        o[ASTRAEA_EN_SIM_TIME] = s.simTime;
    }
    return ASTRAEA_OK;
}

int or_counters(or_ctx* c, uint64_t* out, int n) {
    if (!c || !out) return ASTRAEA_E_INVALID;
    uint64_t v[6] = {0, 0, 0, 0, 0, 0};
    for (auto& s : c->seg) { v[0] += s.interactions; v[1] += s.overflows; v[4] += s.pairs1; v[5] += s.pairs2; }
    for (int k = 0; k < n && k < 6; ++k) out[k] = v[k];
    return ASTRAEA_OK;
}

// Neighbour-table size by brute force over all pairs of each segment, in the FP32 arithmetic of
// the table build: j != i is a candidate of i when dist2f(x_i - x_j) < max(Hb_i, Hb_j)^2 with
// Hb = supportBoundF(h, h0) of the FP32 state.  out[0] = candidates of all particles,
// out[1] = the largest count of one particle.
int or_neighbour_stats(or_ctx* c, uint64_t* out) {
    if (!c || !out) return ASTRAEA_E_INVALID;
    uint64_t total = 0, mx = 0;
    for (auto& s : c->seg) {
        const uint32_t n = s.n;
        std::vector<float> xf(n), yf(n), zf(n), hb(n);
        for (uint32_t k = 0; k < n; ++k) {
            xf[k] = (float)s.x[k]; yf[k] = (float)s.y[k]; zf[k] = (float)s.z[k];
            hb[k] = supportBoundF((float)s.h[k], (float)s.h0[k]);
        }
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total) reduction(max : mx)
        for (uint32_t a = 0; a < n; ++a) {
            uint64_t k = 0;
            for (uint32_t b = 0; b < n; ++b) {
                if (b == a) continue;
                const float r = std::fmax(hb[a], hb[b]);
                if (dist2f(xf[a] - xf[b], yf[a] - yf[b], zf[a] - zf[b]) < r * r) ++k;
            }
            total += k;
            mx = std::max(mx, k);
        }
    }
    out[0] = total;
    out[1] = mx;
    return ASTRAEA_OK;
}

int or_download_tree(or_ctx* c, uint32_t seg, uint64_t* keys, uint32_t* perm, int32_t* left, int32_t* right,
                     int32_t* pI, int32_t* pL, float* cm, float* mn, float* mx) {
    Seg* s = c ? segOf(c, seg) : nullptr;
    if (!s || !s->tree) return ASTRAEA_E_STATE;
    const size_t n = s->n, ni = n >= 2 ? n - 1 : 0;
    if (keys) std::copy(s->keys.begin(), s->keys.end(), keys);
    if (perm) std::copy(s->perm.begin(), s->perm.end(), perm);
    if (left) std::copy(s->left.begin(), s->left.begin() + ni, left);
    if (right) std::copy(s->right.begin(), s->right.begin() + ni, right);
    if (pI) std::copy(s->parI.begin(), s->parI.begin() + ni, pI);
    if (pL) std::copy(s->parL.begin(), s->parL.begin() + n, pL);
    if (cm) std::copy(s->cm4.begin(), s->cm4.end(), cm);
    if (mn) std::copy(s->mn4.begin(), s->mn4.end(), mn);
    if (mx) std::copy(s->mx4.begin(), s->mx4.end(), mx);
    return ASTRAEA_OK;
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

double or_cold_energy_psu(int mat, double rho) { return coldEnergy(mat, rho); }

void or_eos_eval(int mat, float rho, float u, float dmg, double gamma, float* p, float* cs, float* t, uint8_t* ph) {
    const EosOut o = eosEval(mat, (double)rho, (double)u, (double)dmg, gamma);
    if (p) *p = (float)o.p;
    if (cs) *cs = (float)o.cs;
    if (t) *t = (float)o.T;
    if (ph) *ph = (uint8_t)o.phase;
}

void or_eos_from_rho_temp_si(int mat, double rho, double T, double* u, double* p) {
    double uu, pp;
    eosFromRhoTemp(mat, rho, T, uu, pp);
    if (u) *u = uu;
    if (p) *p = pp;
}

double or_invert_density_si(int mat, double p, double T, double guess) { return invertDensity(mat, p, T, guess); }

double or_material_param(int mat, int k) {
    const Mat& p = matOf(mat);
    const double v[19] = {p.rho0, p.a, p.b, p.A, p.B, p.E0, p.Eiv, p.Ecv, p.alpha, p.beta,
                          p.cv, p.tMelt, p.tRef, p.mu, p.Y0, p.epsMin, p.kH2, p.gamma, p.bulk};
    return (k >= 0 && k < 19) ? v[k] : 0.0;
}

// Largest-remnant classification of the telemetry on one body pair's arrays, in input order:
// Simulation::computeTelemetryFromDevice, engine/simulation.cpp:1000-1013 (body sums),
// 1036-1039 (seed = first particle of lowest finite phi), 1094-1112 (spin of body 0 within
// 1.5 target radii), 1115-1178 (five remnant iterations), 1180-1227 (bound planet / disk /
// ejecta).  out: 36 doubles in the order of the ASTRAEA_CL_* indices.
void or_classify_arrays(uint32_t n, const double* x, const double* y, const double* z, const double* vx,
                        const double* vy, const double* vz, const double* m, const float* phi, const uint8_t* body,
                        const float* damage, double targetRadius, double* out) {
    std::fill(out, out + ASTRAEA_CLASS_STRIDE, 0.0);
    double m0 = 0, m1 = 0, s0[6] = {}, s1[6] = {};
    uint32_t seed = 0;
    double minPhi = INFINITY;
    for (uint32_t i = 0; i < n; ++i) {
        const double v[6] = {x[i], y[i], z[i], vx[i], vy[i], vz[i]};
        double* acc = body[i] == 0 ? s0 : s1;
        (body[i] == 0 ? m0 : m1) += m[i];
        for (int k = 0; k < 6; ++k) acc[k] += m[i] * v[k];
        const double ph = (double)phi[i];
        if (std::isfinite(ph) && ph < minPhi) {
            minPhi = ph;
            seed = i;
        }
    }
    for (int k = 0; k < 6; ++k) {
        if (m0 > 0) s0[k] /= m0;
        if (m1 > 0) s1[k] /= m1;
    }
    out[ASTRAEA_CL_BODY0_MASS] = m0;
    out[ASTRAEA_CL_BODY1_MASS] = m1;
    for (int k = 0; k < 6; ++k) {
        out[ASTRAEA_CL_BODY0_COM + k] = s0[k];
        out[ASTRAEA_CL_BODY1_COM + k] = s1[k];
    }
    if (n == 0) return;
    double rem[7] = {x[seed], y[seed], z[seed], vx[seed], vy[seed], vz[seed], 0.0};
    for (int k = 0; k < 6; ++k) out[ASTRAEA_CL_SEED_X + k] = rem[k];
    double remMass = std::max(m0, 0.5 * (m0 + m1));
    double remRadius = std::max(targetRadius, 0.25);
    int iters = 0;
    for (int it = 0; it < 5; ++it) {
        double nm = 0, nx[6] = {}, mr2 = 0;
        const double rSoft = std::max(0.5 * remRadius, 0.1);
        for (uint32_t i = 0; i < n; ++i) {
            const double dx = x[i] - rem[0], dy = y[i] - rem[1], dz = z[i] - rem[2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double dvx = vx[i] - rem[3], dvy = vy[i] - rem[4], dvz = vz[i] - rem[5];
            const double e = 0.5 * (dvx * dvx + dvy * dvy + dvz * dvz) - kG * remMass / std::max(std::sqrt(r2), rSoft);
            if (e < 0.0) {
                const double v[6] = {x[i], y[i], z[i], vx[i], vy[i], vz[i]};
                nm += m[i];
                for (int k = 0; k < 6; ++k) nx[k] += m[i] * v[k];
                mr2 += m[i] * r2;
            }
        }
        if (!(nm > 0.0)) break;
        remMass = nm;
        for (int k = 0; k < 6; ++k) rem[k] = nx[k] / nm;
        remRadius = std::max(0.5 * targetRadius, std::sqrt(5.0 / 3.0) * std::sqrt(mr2 / nm));
        ++iters;
    }
    out[ASTRAEA_CL_REM_MASS] = remMass;
    for (int k = 0; k < 3; ++k) {
        out[ASTRAEA_CL_REM_POS + k] = rem[k];
        out[ASTRAEA_CL_REM_VEL + k] = rem[3 + k];
    }
    out[ASTRAEA_CL_REM_RADIUS] = remRadius;
    out[ASTRAEA_CL_ITERATIONS] = iters;
    const double rSoft = std::max(0.5 * remRadius, 0.1), mu = kG * remMass;
    const double lim2 = (1.5 * targetRadius) * (1.5 * targetRadius);
    for (uint32_t i = 0; i < n; ++i) {
        out[ASTRAEA_CL_MAX_DAMAGE] = std::max(out[ASTRAEA_CL_MAX_DAMAGE], (double)damage[i]);
        const double dx = x[i] - rem[0], dy = y[i] - rem[1], dz = z[i] - rem[2];
        const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
        const double dvx = vx[i] - rem[3], dvy = vy[i] - rem[4], dvz = vz[i] - rem[5];
        const double e = 0.5 * (dvx * dvx + dvy * dvy + dvz * dvz) - mu / std::max(r, rSoft);
        if (e >= 0.0) {
            out[ASTRAEA_CL_EJECTA] += m[i];
        } else {
            if (body[i] != 0) out[ASTRAEA_CL_IMPACTOR_BOUND] += m[i];
            bool disk = false;
            if (r > remRadius && mu > 0.0) {
                const double jx = dy * dvz - dz * dvy, jy = dz * dvx - dx * dvz, jz = dx * dvy - dy * dvx;
                const double j2 = jx * jx + jy * jy + jz * jz;
                const double ecc = std::sqrt(std::max(0.0, 1.0 + 2.0 * e * j2 / (mu * mu)));
                disk = j2 / (mu * (1.0 + ecc)) > remRadius;
            }
            out[disk ? ASTRAEA_CL_DISK : ASTRAEA_CL_BOUND_PLANET] += m[i];
        }
        if (body[i] == 0) {
            const double ex = x[i] - s0[0], ey = y[i] - s0[1], ez = z[i] - s0[2];
            if (ex * ex + ey * ey + ez * ez <= lim2) {
                out[ASTRAEA_CL_SPIN_L] += m[i] * (ez * (vx[i] - s0[3]) - ex * (vz[i] - s0[5]));
                out[ASTRAEA_CL_SPIN_I] += m[i] * (ex * ex + ez * ez);
            }
        }
    }
}

uint64_t or_morton_key(float x, float y, float z, const float* b) {
    const float inv[3] = {invSpanOf(b[0], b[3]), invSpanOf(b[1], b[4]), invSpanOf(b[2], b[5])};
    return mortonKey(x, y, z, b, inv);
}

uint64_t or_morton_key64(double x, double y, double z, const double* b) {
    const double inv[3] = {invSpan64Of(b[0], b[3]), invSpan64Of(b[1], b[4]), invSpan64Of(b[2], b[5])};
    return mortonKey64(x, y, z, b, inv);
}

}  // extern "C"
