// simulation.cpp — astraea::Simulation (include/astraea/simulation.hpp) over the gfx950 C ABI.
//
// Same public behaviour as the reference's engine/simulation.cpp where the hot path is
// concerned: loadScenario splits particles by mass (simulation.cpp:1313-1324), places the two
// bodies on the Keplerian approach orbit (1389-1436), evaluates the initial forces (525-536) and
// records E_0; step() clamps dt_max by the time scale (1716-1717), runs the KDK step (1721-1732)
// on the device, times it with HIP events (1719-1739), integrates the radiated energy (1741-1744)
// and returns telemetry from FP64 device reductions (985-1087).  Errors surface as the
// reference's exceptions: std::invalid_argument for bad configuration, std::runtime_error for
// device failures and call order.
//
// The largest-remnant classification (bound / disk / ejecta mass fractions, living-body centre
// and spin, impactor centre) comes from the device reductions of astraea_classify, and the
// collision outcome from the same rules as simulation.cpp:1229-1281.
//
// Bodies come from the lattice builder (astraea_build_body: uniform density per layer, the
// archetype's layer mass fractions, hydrostatic u) instead of planet_builder.cpp, relaxation is
// not run, and render frames and tracers are not produced.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iomanip>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/astraea/presets.hpp"
#include "../../include/astraea/simulation.hpp"
#include "../../include/astraea_hip.h"

namespace astraea {

namespace {

constexpr double kPi = 3.14159265358979323846;


double rho0Of(MaterialId m) {
    switch (m) {
        case MaterialId::Iron: return 7800.0;
        case MaterialId::WaterIce: return 1000.0;
        case MaterialId::HydrogenHelium: return 0.178;
        case MaterialId::Regolith: return 2000.0;
        default: return 2680.0;
    }
}

void check(int rc, const char* what) {
    if (rc == ASTRAEA_E_INVALID) throw std::invalid_argument(std::string(what) + ": " + astraea_last_error());
    if (rc != ASTRAEA_OK) throw std::runtime_error(std::string(what) + ": " + astraea_last_error());
}

const char* outcomeToString(CollisionOutcome o) {
    switch (o) {
        case CollisionOutcome::Approaching: return "Approaching";
        case CollisionOutcome::Merging: return "Merging";
        case CollisionOutcome::GrazeAndMerge: return "GrazeAndMerge";
        case CollisionOutcome::HitAndRun: return "HitAndRun";
        case CollisionOutcome::CatastrophicDisruption: return "CatastrophicDisruption";
    }
    return "Approaching";
}

double finiteOr0(double v) { return std::isfinite(v) ? v : 0.0; }

void jsonFields(std::ostringstream& os, const TelemetrySnapshot& s) {
    os << std::setprecision(15) << "\"frameSeq\":" << s.frameSeq << ",\"simTimeSec\":" << finiteOr0(s.simTimeSec)
       << ",\"dtSec\":" << finiteOr0(s.dtSec) << ",\"outcome\":\"" << outcomeToString(s.outcome) << "\""
       << ",\"kineticEnergyJ\":" << finiteOr0(s.kineticEnergyJ) << ",\"potentialEnergyJ\":" << finiteOr0(s.potentialEnergyJ)
       << ",\"internalEnergyJ\":" << finiteOr0(s.internalEnergyJ) << ",\"radiatedEnergyJ\":" << finiteOr0(s.radiatedEnergyJ)
       << ",\"totalEnergyJ\":" << finiteOr0(s.totalEnergyJ) << ",\"initialTotalEnergyJ\":" << finiteOr0(s.initialTotalEnergyJ)
       << ",\"energyDriftRel\":" << finiteOr0(s.energyDriftRel) << ",\"momentumKgMS\":[" << finiteOr0(s.momentumKgMS[0])
       << "," << finiteOr0(s.momentumKgMS[1]) << "," << finiteOr0(s.momentumKgMS[2]) << "],\"angularMomentumKgM2S\":["
       << finiteOr0(s.angularMomentumKgM2S[0]) << "," << finiteOr0(s.angularMomentumKgM2S[1]) << ","
       << finiteOr0(s.angularMomentumKgM2S[2]) << "],\"virialRatio\":" << finiteOr0(s.virialRatio)
       << ",\"maxTempK\":" << finiteOr0(s.maxTempK) << ",\"maxPressureGPa\":" << finiteOr0(s.maxPressureGPa)
       << ",\"boundMassFrac\":" << finiteOr0(s.boundMassFrac) << ",\"diskMassFrac\":" << finiteOr0(s.diskMassFrac)
       << ",\"ejectaMassFrac\":" << finiteOr0(s.ejectaMassFrac) << ",\"qRatioQrd\":" << finiteOr0(s.qRatioQrd)
       << ",\"gpuStepMs\":" << finiteOr0(s.gpuStepMs) << ",\"livingBodyComKm\":[" << finiteOr0(s.livingBodyComKm[0]) << ","
       << finiteOr0(s.livingBodyComKm[1]) << "," << finiteOr0(s.livingBodyComKm[2]) << "],\"livingBodyRadiusKm\":"
       << finiteOr0(s.livingBodyRadiusKm) << ",\"livingBodySpinRadS\":" << finiteOr0(s.livingBodySpinRadS)
       << ",\"impactorComKm\":[" << finiteOr0(s.impactorComKm[0]) << "," << finiteOr0(s.impactorComKm[1]) << ","
       << finiteOr0(s.impactorComKm[2]) << "],\"particleCount\":" << s.particleCount << ",\"tracerCount\":" << s.tracerCount
       << ",\"tracerGridRes\":" << s.tracerGridRes << ",\"livingBodyPresent\":" << (s.livingBodyPresent ? "true" : "false");
}

void csvRow(std::ostringstream& os, const TelemetrySnapshot& s) {
    const double v[] = {s.simTimeSec, (double)s.dtSec};
    os << std::setprecision(15) << s.frameSeq << ',' << finiteOr0(v[0]) << ',' << finiteOr0(v[1]) << ','
       << outcomeToString(s.outcome);
    const double rest[] = {s.kineticEnergyJ, s.potentialEnergyJ, s.internalEnergyJ, s.radiatedEnergyJ, s.totalEnergyJ,
                           s.initialTotalEnergyJ, s.energyDriftRel, s.momentumKgMS[0], s.momentumKgMS[1],
                           s.momentumKgMS[2], s.angularMomentumKgM2S[0], s.angularMomentumKgM2S[1],
                           s.angularMomentumKgM2S[2], s.virialRatio, s.maxTempK, s.maxPressureGPa, s.boundMassFrac,
                           s.diskMassFrac, s.ejectaMassFrac, s.qRatioQrd, s.gpuStepMs, s.livingBodyComKm[0],
                           s.livingBodyComKm[1], s.livingBodyComKm[2], s.livingBodyRadiusKm, s.livingBodySpinRadS,
                           s.impactorComKm[0], s.impactorComKm[1], s.impactorComKm[2]};
    for (double d : rest) os << ',' << finiteOr0(d);
    os << ',' << s.particleCount << ',' << s.tracerCount << ',' << s.tracerGridRes << ',' << (s.livingBodyPresent ? 1 : 0)
       << '\n';
}

struct Body {
    std::vector<float> x, y, z, m, h, rho, u, eps;
    std::vector<uint8_t> mat;
    double radiusPsu = 0.0, massKg = 0.0;
};

Body buildBody(const BodyInitSpec& spec, uint32_t n, uint64_t seed) {
    const ArchetypeComposition& comp = getArchetypeComposition(spec.archetype);
    astraea_body_spec bs;
    astraea_default_body_spec(&bs);
    double vPerKg = 0.0;
    bs.num_layers = 0;
    for (uint32_t k = 0; k < comp.layerCount; ++k) {
        const PlanetLayerDef& l = comp.layers[k];
        if (l.massFraction <= 0.0) continue;
        if (l.material == MaterialId::HydrogenHelium)
            throw std::invalid_argument("loadScenario: a body of a given mass with an H/He envelope is not built here "
                                        "(the polytrope builder fixes the radius, not the mass)");
        bs.layer_material[bs.num_layers] = (uint8_t)l.material;
        bs.layer_mass_frac[bs.num_layers] = l.massFraction;
        bs.layer_density_kgm3[bs.num_layers] = 0.0;
        ++bs.num_layers;
        vPerKg += l.massFraction / rho0Of(l.material);
    }
    const double volume = std::max(spec.massKg, 1.0) * vPerKg;
    bs.radius_m = std::cbrt(3.0 * volume / (4.0 * kPi));
    bs.surface_temp_k = spec.surfaceTempK;
    bs.strength_flaws = 1;
    bs.seed = seed;
    Body b;
    for (auto* v : {&b.x, &b.y, &b.z, &b.m, &b.h, &b.rho, &b.u, &b.eps}) v->resize(n);
    b.mat.resize(n);
    astraea_body_out out{b.x.data(), b.y.data(), b.z.data(), b.m.data(), b.h.data(), b.rho.data(), b.u.data(),
                         b.eps.data(), b.mat.data()};
    double mass = 0.0, spacing = 0.0;
    check(astraea_build_body(&bs, n, &out, &mass, &spacing), "astraea_build_body");
    b.radiusPsu = bs.radius_m / units::L_UNIT_M;
    b.massKg = mass;
    return b;
}

// Leinhardt & Stewart (2012) Q/Q*_RD as in engine/simulation.cpp:249-296.
double qRatio(double mT, double mI, double rT, double rI, double vC, double b) {
    if (mI <= 0.0 || mT <= 0.0) return 0.0;
    const double M = mT + mI, mu = mT * mI / M;
    const double qImp = 0.5 * (mu / M) * vC * vC;
    const double rho1 = 1000.0, cStar = 1.9, muBar = 0.36;
    const double rc1 = std::cbrt(3.0 * M / (4.0 * kPi * rho1));
    const double q1 = cStar * 0.8 * kPi * rho1 * units::G_SI * rc1 * rc1;
    const double g = std::max(std::min(mT, mI) / std::max(mT, mI), 1.0e-9);
    const double qHead = q1 * std::pow((g + 1.0) * (g + 1.0) / (4.0 * g), 2.0 / (3.0 * muBar) - 1.0);
    const double bb = std::clamp(std::abs(b), 0.0, 1.0);
    const double RT = std::max(rT, 1.0), RI = std::max(rI, 1.0);
    const double l = (RT + RI) * (1.0 - bb);
    double alpha = 1.0;
    if (l <= 0.0) alpha = 0.02;
    else if (l < 2.0 * RI) alpha = std::clamp(l * l * (3.0 * RI - l) / (4.0 * RI * RI * RI), 0.02, 1.0);
    const double muA = alpha * mI * mT / (mT + alpha * mI);
    const double qObl = qHead * std::pow(std::max(mu / std::max(muA, 1.0e-12 * M), 1.0), 2.0 - 1.5 * muBar);
    return qObl > 0.0 ? qImp / qObl : 0.0;
}

}  // namespace

struct Simulation::Impl {
    uint32_t maxParticles = 0;
    astraea_ctx* ctx = nullptr;
    hipEvent_t evStart = nullptr, evStop = nullptr;
    CollisionScenarioConfig config{};
    SlicePlaneConfig slice{};
    double timeScale = 1.0;
    uint64_t frameSeq = 0;
    double simTimeSec = 0.0;
    float lastDtSec = 0.0f;
    double initialTotalEnergyJ = 0.0;
    double lastGpuStepMs = 0.0;
    double qRatioQrd = 0.0;
    double targetRadiusPsu = 1.0;
    double impactorRadiusPsu = 0.0;
    bool hasImpacted = false;
    uint32_t count = 0;
    bool livingBodyPresent = false;
    TelemetrySnapshot lastSnapshot{};
    std::vector<TelemetrySnapshot> history;

    explicit Impl(uint32_t maxP) : maxParticles(std::max(maxP, 1u)) {
        check(astraea_create(&ctx, 0, maxParticles, 1), "astraea_create");
        if (hipEventCreate(&evStart) != hipSuccess || hipEventCreate(&evStop) != hipSuccess)
            throw std::runtime_error("hipEventCreate failed");
    }
    ~Impl() {
        if (evStart) (void)hipEventDestroy(evStart);
        if (evStop) (void)hipEventDestroy(evStop);
        if (ctx) astraea_destroy(ctx);
    }

    void ensureCapacity(uint32_t n) {
        if (n <= maxParticles) return;
        astraea_destroy(ctx);
        ctx = nullptr;
        maxParticles = n;
        check(astraea_create(&ctx, 0, maxParticles, 1), "astraea_create");
    }

    TelemetrySnapshot telemetry() {
        TelemetrySnapshot s;
        s.frameSeq = frameSeq;
        s.simTimeSec = simTimeSec;
        s.dtSec = lastDtSec;
        s.gpuStepMs = lastGpuStepMs;
        s.particleCount = count;
        s.qRatioQrd = qRatioQrd;
        s.livingBodyPresent = livingBodyPresent;
        s.initialTotalEnergyJ = initialTotalEnergyJ;
        if (count == 0) return s;
        double e[ASTRAEA_ENERGY_STRIDE];
        check(astraea_energy(ctx, e), "astraea_energy");
        s.kineticEnergyJ = e[ASTRAEA_EN_KINETIC] * units::E_UNIT_J;
        s.potentialEnergyJ = e[ASTRAEA_EN_POTENTIAL] * units::E_UNIT_J;
        s.internalEnergyJ = e[ASTRAEA_EN_INTERNAL] * units::E_UNIT_J;
        s.radiatedEnergyJ = e[ASTRAEA_EN_RADIATED] * units::E_UNIT_J;
        s.totalEnergyJ = s.kineticEnergyJ + s.potentialEnergyJ + s.internalEnergyJ;
        if (frameSeq == 0 && simTimeSec == 0.0) {
            initialTotalEnergyJ = s.totalEnergyJ;
            s.initialTotalEnergyJ = initialTotalEnergyJ;
            s.energyDriftRel = 0.0;
        } else {
            s.energyDriftRel = (s.totalEnergyJ + s.radiatedEnergyJ - initialTotalEnergyJ) /
                               std::max(std::abs(initialTotalEnergyJ), 1.0);
        }
        const double mom = units::M_UNIT_KG * units::V_UNIT_MS;
        for (int a = 0; a < 3; ++a) {
            s.momentumKgMS[a] = e[ASTRAEA_EN_MOM_X + a] * mom;
            s.angularMomentumKgM2S[a] = e[ASTRAEA_EN_ANG_X + a] * units::L_ANG_UNIT;
        }
        s.virialRatio = std::abs(s.potentialEnergyJ) > 1.0 ? -2.0 * s.kineticEnergyJ / s.potentialEnergyJ : 0.0;
        s.maxTempK = e[ASTRAEA_EN_MAX_TEMP];
        s.maxPressureGPa = e[ASTRAEA_EN_MAX_PRESSURE] * 1000.0;
        s.livingBodyRadiusKm = livingBodyPresent ? targetRadiusPsu * 1000.0 : 0.0;
        classify(s, e);
        lastSnapshot = s;
        return s;
    }

    // Bound-mass classification and collision outcome (simulation.cpp:1043-1060, 1087-1118,
    // 1229-1281) from the device reductions.
    void classify(TelemetrySnapshot& s, const double* e) {
        double c[ASTRAEA_CLASS_STRIDE];
        check(astraea_classify(ctx, &targetRadiusPsu, c), "astraea_classify");
        const double m0 = c[ASTRAEA_CL_BODY0_MASS], m1 = c[ASTRAEA_CL_BODY1_MASS];
        const double total = e[ASTRAEA_EN_MASS];
        const double* sys = e + ASTRAEA_EN_COM_X;
        const double* com0 = c + ASTRAEA_CL_BODY0_COM;
        const double* com1 = c + ASTRAEA_CL_BODY1_COM;
        if (livingBodyPresent) {
            for (int a = 0; a < 3; ++a) s.livingBodyComKm[a] = ((m0 > 0.0 ? com0[a] : sys[a]) - sys[a]) * 1000.0;
            const double iyy = c[ASTRAEA_CL_SPIN_I];
            s.livingBodySpinRadS = iyy > 0.0 ? c[ASTRAEA_CL_SPIN_L] / iyy / units::T_UNIT_S : config.target.spinOmegaRadS;
        }
        if (m1 > 0.0)
            for (int a = 0; a < 3; ++a) s.impactorComKm[a] = (com1[a] - sys[a]) * 1000.0;
        if (total > 0.0) {
            s.boundMassFrac = c[ASTRAEA_CL_BOUND_PLANET] / total;
            s.diskMassFrac = c[ASTRAEA_CL_DISK] / total;
            s.ejectaMassFrac = c[ASTRAEA_CL_EJECTA] / total;
        }
        if (m1 <= 0.0) {
            s.outcome = CollisionOutcome::Merging;
            return;
        }
        double r[3], v[3];
        for (int a = 0; a < 3; ++a) {
            r[a] = com1[a] - com0[a];
            v[a] = c[ASTRAEA_CL_BODY1_VEL + a] - c[ASTRAEA_CL_BODY0_VEL + a];
        }
        const double dSep = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        const double rDotV = r[0] * v[0] + r[1] * v[1] + r[2] * v[2];
        const double contact = targetRadiusPsu + impactorRadiusPsu;
        if (dSep <= 1.05 * contact || c[ASTRAEA_CL_MAX_DAMAGE] > 0.25 || rDotV >= 0.0) hasImpacted = true;
        if (!hasImpacted && dSep > 1.05 * contact && rDotV < 0.0) {
            s.outcome = CollisionOutcome::Approaching;
            return;
        }
        const double impactorBoundFrac = c[ASTRAEA_CL_IMPACTOR_BOUND] / m1;
        const double bAbs = std::abs(config.impactParamB);
        const double twoBodyE = 0.5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) -
                                units::G_PSU * (m0 + m1) / std::max(dSep, 1.0e-3);
        if (s.boundMassFrac < 0.50 || (qRatioQrd >= 1.0 && (s.ejectaMassFrac >= 0.15 || s.boundMassFrac < 0.72)))
            s.outcome = CollisionOutcome::CatastrophicDisruption;
        else if (rDotV > 0.0 && dSep > 1.05 * contact && (impactorBoundFrac < 0.50 || twoBodyE > 0.0))
            s.outcome = CollisionOutcome::HitAndRun;
        else if (bAbs >= 0.50 && config.impactVelocityRatioVesc > 1.25 && impactorBoundFrac < 0.45)
            s.outcome = CollisionOutcome::HitAndRun;
        else if (bAbs >= 0.45)
            s.outcome = CollisionOutcome::GrazeAndMerge;
        else
            s.outcome = CollisionOutcome::Merging;
    }
};

Simulation::Simulation(uint32_t maxParticles) : impl_(std::make_unique<Impl>(maxParticles)) {}
Simulation::~Simulation() = default;
Simulation::Simulation(Simulation&&) noexcept = default;
Simulation& Simulation::operator=(Simulation&&) noexcept = default;

void Simulation::loadScenario(const CollisionScenarioConfig& config) {
    if (config.totalParticles == 0u) throw std::invalid_argument("Simulation::loadScenario requires totalParticles > 0");
    Impl& I = *impl_;
    I.ensureCapacity(config.totalParticles);
    I.config = config;
    I.frameSeq = 0;
    I.simTimeSec = 0.0;
    I.lastDtSec = 0.0f;
    I.initialTotalEnergyJ = 0.0;
    I.lastGpuStepMs = 0.0;
    I.history.clear();
    I.hasImpacted = false;
    I.impactorRadiusPsu = 0.0;

    // particle split proportional to mass (simulation.cpp:1313-1324)
    const bool hasImpactor = config.impactor.massKg > 0.0 && config.totalParticles >= 2u;
    uint32_t nT = config.totalParticles, nI = 0;
    if (hasImpactor) {
        const double mTot = config.target.massKg + config.impactor.massKg;
        uint32_t k = (uint32_t)std::llround((double)config.totalParticles * (config.impactor.massKg / mTot));
        const uint32_t minK = std::min(config.totalParticles / 2u, 16u);
        k = std::clamp(k, minK, config.totalParticles - 1u);
        nI = k;
        nT = config.totalParticles - k;
    }
    Body t = buildBody(config.target, nT, config.randomSeed);
    I.targetRadiusPsu = t.radiusPsu;
    I.livingBodyPresent = config.target.archetype == PlanetArchetype::LivingTerra;
    const uint32_t n = nT + nI;
    std::vector<float> x(t.x), y(t.y), z(t.z), m(t.m), h(t.h), rho(t.rho), u(t.u), eps(t.eps);
    std::vector<uint8_t> mat(t.mat), body(nT, 0);
    std::vector<float> vx(n, 0.0f), vy(n, 0.0f), vz(n, 0.0f);
    // rigid spin about +Y (simulation.cpp:331-341)
    const double omT = config.target.spinOmegaRadS * units::T_UNIT_S;
    for (uint32_t i = 0; i < nT; ++i) {
        vx[i] = (float)(omT * (double)z[i]);
        vz[i] = (float)(-omT * (double)x[i]);
    }
    I.qRatioQrd = 0.0;
    if (hasImpactor) {
        Body b = buildBody(config.impactor, nI, config.randomSeed + 1ULL);
        I.impactorRadiusPsu = b.radiusPsu;
        const double omI = config.impactor.spinOmegaRadS * units::T_UNIT_S;
        for (uint32_t i = 0; i < nI; ++i) {
            x.push_back(b.x[i]); y.push_back(b.y[i]); z.push_back(b.z[i]); m.push_back(b.m[i]);
            h.push_back(b.h[i]); rho.push_back(b.rho[i]); u.push_back(b.u[i]); eps.push_back(b.eps[i]);
            mat.push_back(b.mat[i]); body.push_back(1);
            vx[nT + i] = (float)(omI * (double)b.z[i]);
            vz[nT + i] = (float)(-omI * (double)b.x[i]);
        }
        double vEscPsu = 0.0;
        check(astraea_place_two_bodies(nT, nI, x.data(), y.data(), z.data(), vx.data(), vy.data(), vz.data(), m.data(),
                                       t.radiusPsu, b.radiusPsu, config.impactVelocityRatioVesc, config.impactParamB,
                                       config.initialSepRadii, &vEscPsu),
              "astraea_place_two_bodies");
        I.qRatioQrd = qRatio(t.massKg, b.massKg, t.radiusPsu * units::L_UNIT_M, b.radiusPsu * units::L_UNIT_M,
                             std::max(0.0, config.impactVelocityRatioVesc) * vEscPsu * units::V_UNIT_MS,
                             config.impactParamB);
    }
    astraea_params p;
    astraea_default_params(&p);
    check(astraea_set_params(I.ctx, &p), "astraea_set_params");
    check(astraea_set_layout(I.ctx, 1, &n), "astraea_set_layout");
    check(astraea_upload_field(I.ctx, ASTRAEA_F_POS_X, x.data()), "upload");
    check(astraea_upload_field(I.ctx, ASTRAEA_F_POS_Y, y.data()), "upload");
    check(astraea_upload_field(I.ctx, ASTRAEA_F_POS_Z, z.data()), "upload");
    check(astraea_upload_field(I.ctx, ASTRAEA_F_VEL_X, vx.data()), "upload");
    check(astraea_upload_field(I.ctx, ASTRAEA_F_VEL_Y, vy.data()), "upload");
    check(astraea_upload_field(I.ctx, ASTRAEA_F_VEL_Z, vz.data()), "upload");
    check(astraea_upload_field(I.ctx, ASTRAEA_F_MASS, m.data()), "upload");
    check(astraea_upload_field(I.ctx, ASTRAEA_F_H, h.data()), "upload");
    check(astraea_upload_field(I.ctx, ASTRAEA_F_H0, h.data()), "upload");
    check(astraea_upload_field(I.ctx, ASTRAEA_F_RHO, rho.data()), "upload");
    check(astraea_upload_field(I.ctx, ASTRAEA_F_U, u.data()), "upload");
    check(astraea_upload_field(I.ctx, ASTRAEA_F_EPS_FAIL, eps.data()), "upload");
    check(astraea_upload_field(I.ctx, ASTRAEA_F_MATERIAL, mat.data()), "upload");
    check(astraea_upload_field(I.ctx, ASTRAEA_F_BODY, body.data()), "upload");
    I.count = n;
    check(astraea_evaluate_initial_forces(I.ctx), "astraea_evaluate_initial_forces");
    check(astraea_synchronize(I.ctx), "astraea_synchronize");
    I.telemetry();
}

void Simulation::relaxBody(uint8_t bodyId, uint32_t steps, double& outFinalVrmsOverVesc) {
    (void)bodyId;
    (void)steps;
    if (impl_->count == 0u) throw std::runtime_error("Simulation::relaxBody called before loadScenario");
    // Relaxation is not part of this build: the state is left as loaded and the history reset
    // as the reference does (simulation.cpp:1464-1477).
    outFinalVrmsOverVesc = 0.0;
    impl_->frameSeq = 0;
    impl_->simTimeSec = 0.0;
    impl_->history.clear();
    impl_->telemetry();
}

TelemetrySnapshot Simulation::step(double maxDtSec) {
    Impl& I = *impl_;
    if (I.count == 0u) throw std::runtime_error("Simulation::step called before loadScenario");
    const double eff = std::max(1.0e-5, maxDtSec * I.timeScale);
    const float maxDtPsu = (float)(eff / units::T_UNIT_S);
    hipStream_t s = (hipStream_t)astraea_stream(I.ctx);
    if (hipEventRecord(I.evStart, s) != hipSuccess) throw std::runtime_error("hipEventRecord failed");
    check(astraea_step(I.ctx, maxDtPsu, 1, 0), "astraea_step");
    if (hipEventRecord(I.evStop, s) != hipSuccess || hipEventSynchronize(I.evStop) != hipSuccess)
        throw std::runtime_error("hipEvent failed");
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, I.evStart, I.evStop);
    I.lastGpuStepMs = ms;
    check(astraea_synchronize(I.ctx), "astraea_synchronize");
    float dt = 0.0f;
    check(astraea_get_dt(I.ctx, &dt), "astraea_get_dt");
    const double dtSec = (double)dt * units::T_UNIT_S;
    I.simTimeSec += dtSec;
    I.lastDtSec = (float)dtSec;
    I.frameSeq += 1;
    TelemetrySnapshot snap = I.telemetry();
    I.history.push_back(snap);
    return snap;
}

TelemetrySnapshot Simulation::stepForDuration(double durationSec, double maxDtSec) {
    if (impl_->count == 0u) throw std::runtime_error("Simulation::stepForDuration called before loadScenario");
    if (durationSec <= 0.0) return computeTelemetry();
    const double end = impl_->simTimeSec + durationSec;
    TelemetrySnapshot snap = impl_->lastSnapshot;
    while (impl_->simTimeSec < end - 1.0e-9) {
        const double remaining = end - impl_->simTimeSec;
        snap = step(std::min(maxDtSec, std::max(1.0e-5, remaining / impl_->timeScale)));
    }
    return snap;
}

void Simulation::setTimeScale(double scale) {
    if (!std::isfinite(scale)) return;
    impl_->timeScale = std::clamp(scale, 0.1, 100.0);
}

void Simulation::setSlicePlane(const SlicePlaneConfig& slice) {
    SlicePlaneConfig s = slice;
    const double len = std::sqrt((double)slice.normalX * slice.normalX + (double)slice.normalY * slice.normalY +
                                 (double)slice.normalZ * slice.normalZ);
    if (len > 1.0e-9) {
        s.normalX = (float)(slice.normalX / len);
        s.normalY = (float)(slice.normalY / len);
        s.normalZ = (float)(slice.normalZ / len);
    } else {
        s.normalX = 0.0f;
        s.normalY = 1.0f;
        s.normalZ = 0.0f;
    }
    impl_->slice = s;
}

TelemetrySnapshot Simulation::computeTelemetry() {
    if (impl_->count == 0u) return impl_->lastSnapshot;
    return impl_->telemetry();
}

void Simulation::extractRenderFrame(std::vector<uint8_t>& outParticleBytes, std::vector<uint8_t>& outTracerBytes,
                                    std::vector<uint8_t>* outSpeedBytes) {
    // Render packing (sph_gravity.cu:1554-1743) is outside the physics path built here.
    outParticleBytes.clear();
    outTracerBytes.clear();
    if (outSpeedBytes) outSpeedBytes->clear();
}

std::string Simulation::exportTelemetryCsv() const {
    std::ostringstream os;
    os << "frameSeq,simTimeSec,dtSec,outcome,kineticEnergyJ,potentialEnergyJ,internalEnergyJ,radiatedEnergyJ,"
          "totalEnergyJ,initialTotalEnergyJ,energyDriftRel,momentumKgMS_x,momentumKgMS_y,momentumKgMS_z,"
          "angularMomentumKgM2S_x,angularMomentumKgM2S_y,angularMomentumKgM2S_z,virialRatio,maxTempK,maxPressureGPa,"
          "boundMassFrac,diskMassFrac,ejectaMassFrac,qRatioQrd,gpuStepMs,livingBodyComKm_x,livingBodyComKm_y,"
          "livingBodyComKm_z,livingBodyRadiusKm,livingBodySpinRadS,impactorComKm_x,impactorComKm_y,impactorComKm_z,"
          "particleCount,tracerCount,tracerGridRes,livingBodyPresent\n";
    if (impl_->history.empty()) csvRow(os, impl_->lastSnapshot);
    for (const TelemetrySnapshot& s : impl_->history) csvRow(os, s);
    return os.str();
}

std::string Simulation::exportTelemetryJson() const {
    std::ostringstream os;
    const TelemetrySnapshot& latest = impl_->history.empty() ? impl_->lastSnapshot : impl_->history.back();
    os << "{\"preset\":\"";
    for (char ch : impl_->config.presetName) {
        if (ch == '"' || ch == '\\') os << '\\';
        os << ch;
    }
    os << "\",";
    jsonFields(os, latest);
    os << ",\"history\":[";
    if (impl_->history.empty()) {
        os << "{";
        jsonFields(os, impl_->lastSnapshot);
        os << "}";
    }
    for (size_t k = 0; k < impl_->history.size(); ++k) {
        os << (k ? ",{" : "{");
        jsonFields(os, impl_->history[k]);
        os << "}";
    }
    os << "]}";
    return os.str();
}

const std::vector<TelemetrySnapshot>& Simulation::telemetryHistory() const { return impl_->history; }
const CollisionScenarioConfig& Simulation::activeConfig() const { return impl_->config; }
const SlicePlaneConfig& Simulation::slicePlane() const { return impl_->slice; }
double Simulation::timeScale() const { return impl_->timeScale; }

}  // namespace astraea
