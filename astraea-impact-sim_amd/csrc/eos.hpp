// eos.hpp — Tillotson and H/He equations of state, FP64, callable on host and device.
//
// Restates engine/eos.cu of the reference: the material table (eos.cu:10-116),
// cold compression energy (283-318), compressed / expanded / two-phase Tillotson
// with analytic dP/drho, dP/du (136-227, 352-378), sound speed with the
// condensed-phase floor (380-387) and the phase classification (229-267).
// MI355X runs FP64 vector math at the same rate as unpacked FP32, so the EOS
// keeps the reference's FP64 internals.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace astraea_hip {

constexpr double kRhoUnit = 1.0e6;  // kg/m^3 per PSU density (engine/units.hpp:20)
constexpr double kPUnit = 1.0e12;   // Pa per PSU pressure (units.hpp:21)
constexpr double kUUnit = 1.0e6;    // J/kg per PSU specific energy (units.hpp:22)

enum Phase : uint8_t { kSolid = 0, kMolten = 1, kVapor = 2, kPlasma = 3, kLiquidWater = 4, kGas = 5 };

struct Material {
    double rho0;      // kg/m^3
    double a, b;      // Tillotson a, b
    double A, B;      // Pa
    double E0, Eiv, Ecv;  // J/kg
    double alpha, beta;
    double cv;        // J/(kg K)
    double tMelt, tRef;   // K
    double mu, Y0;    // shear modulus, yield (Pa)
    double epsMin;    // baseline Weibull failure strain
    double kH2;       // H/He polytrope constant (Pa m^6/kg^2)
    double bulk;      // bulk modulus of the strength table (Pa), sph_gravity.cu:42-73
};

// Material index -> parameters; indices > 4 fall back to granite (eos.cu:270-279).
__host__ __device__ inline Material material(int m) {
    switch (m) {
        case 0: return {7800.0, 0.5, 1.5, 1.28e11, 1.05e11, 9.50e6, 2.40e6, 8.67e6, 5.0, 5.0, 449.0, 1811.0, 293.0, 7.70e10, 6.80e8, 1.0e-3, 0.0, 1.28e11};
        case 2: return {1000.0, 0.7, 0.15, 2.18e9, 1.325e10, 7.00e6, 4.19e5, 2.68e6, 10.0, 5.0, 4184.0, 273.15, 273.15, 3.50e9, 1.00e7, 5.0e-4, 0.0, 2.18e9};
        case 3: return {0.178, 0.4, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 8260.0, 0.0, 0.0, 0.0, 0.0, 1.0, 2.08e5, 0.0};
        case 4: return {2000.0, 0.5, 1.3, 1.20e10, 1.20e10, 1.60e7, 3.50e6, 1.80e7, 5.0, 5.0, 800.0, 1500.0, 293.0, 8.00e9, 5.00e7, 5.0e-5, 0.0, 1.20e10};
        default: return {2680.0, 0.5, 1.3, 1.80e10, 1.80e10, 1.60e7, 3.50e6, 1.80e7, 5.0, 5.0, 790.0, 1600.0, 293.0, 2.20e10, 3.50e9, 1.0e-4, 0.0, 1.80e10};
    }
}

__host__ __device__ inline int materialIndex(int m) { return (m >= 0 && m <= 4) ? m : 1; }
__host__ __device__ inline bool isHHe(int m) { return materialIndex(m) == 3; }

__host__ __device__ inline double dmax(double a, double b) { return a > b ? a : b; }
__host__ __device__ inline double dmin(double a, double b) { return a < b ? a : b; }
__host__ __device__ inline double dclamp(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Cold compression energy u_cold(rho) in PSU: the integral of P_cold/rho^2 from rho0,
// with a 5-term series below mu = 1e-5 (eos.cu:283-318).
__host__ __device__ inline double coldEnergyPsu(int mat, double rhoPsu) {
    const Material p = material(mat);
    const double rho = dmax(rhoPsu, 0.0);
    if (isHHe(mat)) return p.kH2 * rho;
    const double rho0 = p.rho0 / kRhoUnit;
    const double eta = rho / rho0;
    if (eta <= 1.0) return 0.0;
    const double mu = eta - 1.0;
    const double A = p.A / kPUnit, B = p.B / kPUnit;
    double ta, tb;
    if (mu < 1.0e-5) {
        const double m2 = mu * mu, m3 = m2 * mu, m4 = m3 * mu, m5 = m4 * mu;
        ta = 0.5 * m2 - (2.0 / 3.0) * m3 + 0.75 * m4 - 0.8 * m5;
        tb = (1.0 / 3.0) * m3 - 0.5 * m4 + 0.6 * m5;
    } else {
        const double l = log1p(mu);
        ta = l - mu / eta;
        tb = mu * (2.0 + mu) / eta - 2.0 * l;
    }
    return (A * ta + B * tb) / rho0;
}

struct PressureState {
    double p, dpdrho, dpdu;
    bool damagedTension;
};

// Compressed Tillotson form; negative pressure is scaled by (1 - D) (eos.cu:136-180).
__host__ __device__ inline PressureState tillotsonCompressed(const Material& m, double rho, double rho0, double eta,
                                                            double mu, double u, double damage) {
    const double A = m.A / kPUnit, B = m.B / kPUnit, E0 = m.E0 / kUUnit;
    const double bq = mu >= 0.0 ? B : 0.0;
    const double den = E0 * eta * eta;
    const double ratio = den > 1.0e-30 ? u / den : 1.0e30;
    const double iw = 1.0 / (ratio + 1.0);
    const double dw = ratio > 1.0e15 ? 0.0 : ratio * iw * iw;
    PressureState s;
    s.p = (m.a + m.b * iw) * rho * u + A * mu + bq * mu * mu;
    s.dpdrho = (m.a + m.b * iw) * u + 2.0 * m.b * u * dw + (A + 2.0 * bq * mu) / rho0;
    s.dpdu = (m.a + m.b * iw * iw) * rho;
    s.damagedTension = false;
    if (s.p < 0.0 && damage > 0.0) {
        const double keep = 1.0 - damage;
        s.p *= keep;
        s.dpdrho *= keep;
        s.dpdu *= keep;
        s.damagedTension = true;
    }
    return s;
}

// Expanded Tillotson form with exp(-alpha x^2), exp(-beta x) cut-offs (eos.cu:182-227).
__host__ __device__ inline PressureState tillotsonExpanded(const Material& m, double rho, double rho0, double eta,
                                                          double mu, double u) {
    const double A = m.A / kPUnit, E0 = m.E0 / kUUnit;
    const double x = 1.0 / eta - 1.0;
    const double ax2 = m.alpha * x * x, bx = m.beta * x;
    const double ea = ax2 < 700.0 ? exp(-ax2) : 0.0;
    PressureState s;
    s.damagedTension = false;
    if (ea == 0.0) {
        s.p = m.a * rho * u;
        s.dpdrho = m.a * u;
        s.dpdu = m.a * rho;
        return s;
    }
    const double eb = bx < 700.0 ? exp(-bx) : 0.0;
    const double den = E0 * eta * eta;
    const double ratio = den > 1.0e-30 ? u / den : 1.0e30;
    const double iw = 1.0 / (ratio + 1.0);
    const double dw = ratio > 1.0e15 ? 0.0 : ratio * iw * iw;
    const double ideal = m.a * rho * u;
    const double inner = m.b * rho * u * iw + A * mu * eb;
    s.p = ideal + inner * ea;
    const double r0r2 = rho0 / (rho * rho);
    const double eDrho = eb > 0.0 ? (A / rho0 + A * mu * m.beta * r0r2) * eb : 0.0;
    s.dpdrho = m.a * u + ea * (m.b * u * iw + 2.0 * m.b * u * dw + eDrho) + 2.0 * m.alpha * x * r0r2 * (s.p - ideal);
    s.dpdu = (m.a + m.b * ea * iw * iw) * rho;
    return s;
}

// Phase from temperature, energy and density (eos.cu:229-267).
__host__ __device__ inline uint8_t classifyPhase(int mat, const Material& m, double rho, double u, double T) {
    const int mi = materialIndex(mat);
    if (mi == 3) return T > 10000.0 ? kPlasma : kGas;
    const double rhoSI = rho * kRhoUnit, ecv = m.Ecv / kUUnit;
    if (mi == 2) {
        if (T > 15000.0) return kPlasma;
        if (u >= ecv || (T > 647.0 && rhoSI < 322.0)) return kVapor;
        if (T >= 273.15) return kLiquidWater;
        return kSolid;
    }
    if (T > 12000.0) return kPlasma;
    if (u >= ecv && rhoSI < 0.8 * m.rho0) return kVapor;
    if (T >= m.tMelt) return kMolten;
    return kSolid;
}

struct EosResult {
    double p, cs, T;
    uint8_t phase;
};

// (rho, u, D) in PSU -> pressure, sound speed, temperature, phase (eos.cu:320-388).
// `gammaHHe` is the H/He ideal-gas gamma (the reference hard-codes 1.4, eos.cu:93).
__host__ __device__ inline EosResult evaluateEos(int mat, double rhoIn, double u, double damageIn, double gammaHHe) {
    const Material m = material(mat);
    const double rho = dmax(rhoIn, 1.0e-12);
    const double dmg = dclamp(damageIn, 0.0, 1.0);
    const double cv = m.cv / kUUnit;
    const double uc = coldEnergyPsu(mat, rho);
    EosResult r;
    if (isHHe(mat)) {
        const double uth = dmax(0.0, u - uc);
        const double T = dmax(3.0, uth / cv);
        r.p = m.kH2 * rho * rho + (gammaHHe - 1.0) * rho * uth;
        const double cs2 = 2.0 * m.kH2 * rho + gammaHHe * (gammaHHe - 1.0) * uth;
        r.cs = dmax(0.05, sqrt(dmax(0.0, cs2)));
        r.T = T;
        r.phase = classifyPhase(mat, m, rho, u, T);
        return r;
    }
    const double rho0 = m.rho0 / kRhoUnit, eta = rho / rho0, mu = eta - 1.0;
    const double T = dmax(3.0, (u - uc) / cv);
    const double uE = uc + dmax(0.0, u - uc - cv * m.tRef);
    const double Eiv = m.Eiv / kUUnit, Ecv = m.Ecv / kUUnit, A = m.A / kPUnit;
    PressureState s;
    bool tension = false;
    if (eta >= 1.0 || uE < Eiv) {
        s = tillotsonCompressed(m, rho, rho0, eta, mu, uE, dmg);
        tension = s.damagedTension;
    } else if (uE >= Ecv) {
        s = tillotsonExpanded(m, rho, rho0, eta, mu, uE);
    } else {
        const PressureState c = tillotsonCompressed(m, rho, rho0, eta, mu, uE, dmg);
        const PressureState e = tillotsonExpanded(m, rho, rho0, eta, mu, uE);
        tension = c.damagedTension;
        const double span = dmax(1.0e-30, Ecv - Eiv);
        const double f = dclamp((uE - Eiv) / span, 0.0, 1.0);
        s.p = (1.0 - f) * c.p + f * e.p;
        s.dpdrho = (1.0 - f) * c.dpdrho + f * e.dpdrho;
        s.dpdu = (1.0 - f) * c.dpdu + f * e.dpdu + (e.p - c.p) / span;
    }
    const double cs2 = s.dpdrho + (s.p / (rho * rho)) * s.dpdu;
    const bool condensed = eta >= 1.0 || (uE < Eiv && rho >= 0.9 * rho0 && !tension);
    const double cmin = condensed ? dmax(0.05, sqrt(A / rho0)) : 0.05;
    r.p = s.p;
    r.cs = dmax(cmin, sqrt(dmax(0.0, cs2)));
    r.T = T;
    r.phase = classifyPhase(mat, m, rho, u, T);
    return r;
}

}  // namespace astraea_hip
